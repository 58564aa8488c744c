// devmem.h — grow-only device buffers owned by a context.
#pragma once

#include "common.h"

#include <unordered_map>

namespace plan {

/* A named, grow-only device allocation pool.  Buffers keep their capacity
 * across calls so repeated pipeline runs on one context do not allocate.
 * Built over a fixed device region (base, cap) instead, it never allocates:
 * each name gets a 256-byte aligned slice carved from the region (reused
 * while it is large enough), and a request that does not fit throws. */
class Arena {
public:
    Arena() {}
    Arena(void *base, size_t cap) : base_(static_cast<char *>(base)), cap_(cap) {}
    ~Arena() { release(); }
    void *get(const std::string &name, size_t bytes)
    {
        if (bytes == 0) bytes = 16;
        auto &b = bufs_[name];
        if (base_) {
            if (b.cap < bytes) {
                const size_t need = (bytes + 255) & ~(size_t)255;
                if (need > cap_ - used_) throw std::runtime_error("fixed scratch region exhausted by " + name);
                b.p = base_ + used_;
                b.cap = need;
                used_ += need;
            }
            return b.p;
        }
        if (b.cap < bytes) {
            if (b.p) HIPCHK(hipFree(b.p));
            b.p = nullptr;
            size_t cap = bytes + bytes / 8 + 256;
            HIPCHK(hipMalloc(&b.p, cap));
            b.cap = cap;
        }
        return b.p;
    }
    template <class T> T *get_t(const std::string &name, size_t count)
    {
        return static_cast<T *>(get(name, count * sizeof(T)));
    }
    void release()
    {
        if (!base_)
            for (auto &kv : bufs_)
                if (kv.second.p) (void)hipFree(kv.second.p);
        bufs_.clear();
        used_ = 0;
    }
    size_t bytes() const
    {
        size_t s = 0;
        for (auto &kv : bufs_) s += kv.second.cap;
        return s;
    }

private:
    struct Buf {
        void *p = nullptr;
        size_t cap = 0;
    };
    std::unordered_map<std::string, Buf> bufs_;
    char *base_ = nullptr;   /* fixed region, or null for hipMalloc'ed buffers */
    size_t cap_ = 0, used_ = 0;
};

/* device bytes owned by one call (freed when it returns) */
struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { HIPCHK(hipMalloc(&p, bytes ? bytes : 16)); }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

/* pinned host staging buffer */
class HostPinned {
public:
    ~HostPinned()
    {
        if (p_) (void)hipHostFree(p_);
    }
    void *get(size_t bytes)
    {
        if (cap_ < bytes) {
            if (p_) HIPCHK(hipHostFree(p_));
            p_ = nullptr;
            HIPCHK(hipHostMalloc(&p_, bytes + 4096, hipHostMallocDefault));
            cap_ = bytes + 4096;
        }
        return p_;
    }

private:
    void *p_ = nullptr;
    size_t cap_ = 0;
};

/* device primitives (scan_sort.hip) */
void scan_exclusive_u32(Arena &a, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t s,
                        uint32_t *total_dev = nullptr);
/* Stable LSD radix sort of 16-byte records {w0,w1,w2,payload} on the 96-bit
 * key (w2:w1:w0).  nwords selects how many key words take part (1..3). The
 * result is left in *keys (buffers may be swapped with *tmp). */
void radix_sort_96(Arena &a, uint4 **keys, uint4 **tmp, uint32_t n, int nwords, hipStream_t s,
                   int first_digit = 0);

}  // namespace plan
