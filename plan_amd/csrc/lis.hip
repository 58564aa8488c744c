// lis.hip — ParticleSet<3>::ReadLisFile (src/tree.hpp:1300-1583) with the
// particle work on the device.  The host only parses the file headers and
// streams the record bytes, unchanged, into HBM through two pinned chunks
// (file reads overlap the DMA).  On the device a kernel stages each block's
// records through LDS and makes the sort key (cpu_id, id_in_run); the radix
// sort puts the records in the reference's order (src/tree.hpp:1552-1557),
// and a second kernel converts each record to the particle set: binary64
// positions, velocities and grid density exactly as static_cast<double> of
// the floats, the species, and id = position after the sort
// (src/tree.hpp:1558-1560).

#include "ctx.h"

#include <cstdio>
#include <cstring>
#include <memory>

namespace plan {

/* One .lis record, little-endian LP64 (src/tree.hpp:1384-1417, 1488-1512):
 * float pos[3], float vel[3], float density, int property_index,
 * long id_in_run, int cpu_id — 44 bytes, 11 words. */
constexpr uint32_t LIS_REC_BYTES = 44, LIS_REC_WORDS = 11;
constexpr int LIS_BLK = 256;

/* Sort keys: the reference compares cpu_id (the record's int cut to 16 bits)
 * and then id_in_run (the record's long cut to 32 bits).  The 48-bit key sits
 * in bits 16-63 of (w1:w0), so the sort starts at byte 2; the payload is the
 * record's position, which keeps equal keys in file order. */
__global__ __launch_bounds__(LIS_BLK) void k_lis_keys(const uint32_t *__restrict__ rec, uint32_t n,
                                                      uint4 *__restrict__ keys)
{
    __shared__ uint32_t w[LIS_BLK * LIS_REC_WORDS];
    const uint32_t b0 = blockIdx.x * LIS_BLK;
    const uint32_t nb = min((uint32_t)LIS_BLK, n - b0);
    const uint32_t *src = rec + (size_t)b0 * LIS_REC_WORDS;
    for (uint32_t k = threadIdx.x; k < nb * LIS_REC_WORDS; k += LIS_BLK) w[k] = src[k];
    __syncthreads();
    const uint32_t t = threadIdx.x;
    if (t >= nb) return;
    const uint32_t idr = w[t * LIS_REC_WORDS + 8]; /* low word of id_in_run */
    const uint32_t cpu = w[t * LIS_REC_WORDS + 10] & 0xffffu;
    keys[b0 + t] = make_uint4(idr << 16, (idr >> 16) | (cpu << 16), 0u, b0 + t);
}

/* Particle j of the set from the record the sort put at position j.  A
 * species outside the header's table sets *bad (the reference would index
 * mass_per_particle out of range). */
__global__ void k_lis_particles(const uint32_t *__restrict__ rec, const uint4 *__restrict__ keys, uint32_t n,
                                uint32_t num_types, double4 *__restrict__ P, double4 *__restrict__ V,
                                int32_t *__restrict__ prop, uint32_t *__restrict__ id, uint32_t *__restrict__ bad)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t *r = rec + (size_t)keys[j].w * LIS_REC_WORDS;
    float f[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) f[k] = __uint_as_float(r[k]);
    const int32_t sp = (int32_t)r[7];
    P[j] = make_double4((double)f[0], (double)f[1], (double)f[2], (double)f[6]);
    V[j] = make_double4((double)f[3], (double)f[4], (double)f[5], 0.0);
    prop[j] = sp;
    id[j] = j;
    if (sp < 0 || (uint32_t)sp >= num_types) atomicOr(bad, 1u);
}

static uint32_t lis_read_u32(const uint32_t *dp, hipStream_t s)
{
    uint32_t v = 0;
    HIPCHK(hipMemcpyAsync(&v, dp, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    return v;
}

bool dev_lis_to_particles(Arena &a, const uint32_t *rec, uint32_t n, uint32_t num_types, DevParticles &ps,
                          hipStream_t s)
{
    DevBuf kb((size_t)n * 2 * sizeof(uint4)); /* keys and the sort's second buffer, for this call only */
    uint4 *k0 = kb.as<uint4>(), *k1 = k0 + n;
    uint32_t *bad = a.get_t<uint32_t>("lis_bad", 1);
    HIPCHK(hipMemsetAsync(bad, 0, 4, s));
    hipLaunchKernelGGL(k_lis_keys, dim3(ceil_div(n, LIS_BLK)), dim3(LIS_BLK), 0, s, rec, n, k0);
    LAUNCHCHK();
    radix_sort_96(a, &k0, &k1, n, 2, s, 2);
    hipLaunchKernelGGL(k_lis_particles, dim3(ceil_div(n, 256)), dim3(256), 0, s, rec, k0, n, num_types, ps.pos,
                       ps.vel, ps.prop, ps.id, bad);
    LAUNCHCHK();
    return lis_read_u32(bad, s) == 0;
}

/* ---- host side: headers and the streamed upload ------------------------------ */

namespace {

struct FileCloser {
    void operator()(FILE *f) const
    {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

template <class T> bool read_pod(FILE *f, T *v, size_t n = 1) { return std::fread(v, sizeof(T), n, f) == n; }

/* Copies byte ranges of files to consecutive device addresses through two
 * pinned chunks: the read of one chunk overlaps the DMA of the other. */
class Uploader {
public:
    static constexpr size_t CHUNK = (size_t)16 << 20;
    explicit Uploader(hipStream_t s) : s_(s)
    {
        for (int b = 0; b < 2; ++b) {
            HIPCHK(hipHostMalloc(&buf_[b], CHUNK, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming | hipEventBlockingSync));
        }
    }
    ~Uploader()
    {
        for (int b = 0; b < 2; ++b) {
            if (ev_[b]) {
                (void)hipEventSynchronize(ev_[b]);
                (void)hipEventDestroy(ev_[b]);
            }
            if (buf_[b]) (void)hipHostFree(buf_[b]);
        }
    }
    /* bytes of f from its current position to dst; false if the file ends first */
    bool copy(FILE *f, uint8_t *dst, size_t bytes)
    {
        while (bytes > 0) {
            const size_t m = bytes < CHUNK ? bytes : CHUNK;
            HIPCHK(hipEventSynchronize(ev_[cur_])); /* the chunk's previous copy is done */
            if (std::fread(buf_[cur_], 1, m, f) != m) return false;
            HIPCHK(hipMemcpyAsync(dst, buf_[cur_], m, hipMemcpyHostToDevice, s_));
            HIPCHK(hipEventRecord(ev_[cur_], s_));
            cur_ ^= 1;
            dst += m;
            bytes -= m;
        }
        return true;
    }

private:
    hipStream_t s_;
    void *buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    int cur_ = 0;
};

}  // namespace

/* Header of the first file (ReadSingleLisFile / ReadMultipleLisFile,
 * src/tree.hpp:1306-1337, 1439-1466) and the particle counts of all files
 * (the others' counts sit after the first file's header length,
 * src/tree.hpp:1344-1350).  Returns an error message or "" on success; io
 * tells a file that cannot be opened or is too short from invalid contents. */
std::string lis_read_headers(const char *const *paths, uint32_t num_paths, LisHeader &h,
                             std::vector<uint64_t> &counts, bool &io)
{
    io = true;
    auto invalid = [&io](std::string m) {
        io = false;
        return m;
    };
    if (!paths || num_paths == 0) return invalid("no .lis file given");
    counts.assign(num_paths, 0);
    uint64_t total = 0;
    for (uint32_t p = 0; p < num_paths; ++p) {
        if (!paths[p]) return invalid("NULL .lis path");
        File f(std::fopen(paths[p], "rb"));
        if (!f) return std::string("Failed to open file ") + paths[p];
        if (p == 0) {
            float cl[12], v;
            int32_t nt;
            if (!read_pod(f.get(), cl, 12) || !read_pod(f.get(), &nt)) return std::string("short header in ") + paths[p];
            for (int k = 0; k < 12; ++k) h.coor_lim[k] = (double)cl[k];
            if (nt < 1 || nt > PLAN_MAX_TYPES) return invalid("num_types in the .lis header must be 1..16");
            h.num_types = (uint32_t)nt;
            for (int k = 0; k < nt; ++k) {
                if (!read_pod(f.get(), &v)) return std::string("short header in ") + paths[p];
                h.type_info[k] = (double)v;
            }
            if (!read_pod(f.get(), &v)) return std::string("short header in ") + paths[p];
            h.time = (double)v;
            if (!read_pod(f.get(), &v)) return std::string("short header in ") + paths[p];
            h.dt = (double)v;
        } else if (std::fseek(f.get(), (long)((14 + h.num_types) * sizeof(float) + sizeof(int32_t)), SEEK_SET) != 0) {
            return std::string("short header in ") + paths[p];
        }
        int64_t np;
        if (!read_pod(f.get(), &np)) return std::string("short header in ") + paths[p];
        if (np < 0) return invalid(std::string("negative particle count in ") + paths[p]);
        counts[p] = (uint64_t)np;
        total += (uint64_t)np;
        if (total > 0xffffffffull) return invalid("more than 2^32 - 1 particles");
    }
    if (total == 0) return invalid("no particles");
    h.body_offset = 68 + 4 * (uint64_t)h.num_types;
    h.num_particles = (uint32_t)total;
    return "";
}

std::string lis_upload_records(const char *const *paths, uint32_t num_paths, const LisHeader &h,
                               const std::vector<uint64_t> &counts, uint8_t *dst, hipStream_t s)
{
    Uploader up(s);
    for (uint32_t p = 0; p < num_paths; ++p) {
        File f(std::fopen(paths[p], "rb"));
        if (!f) return std::string("Failed to open file ") + paths[p];
        if (std::fseek(f.get(), (long)h.body_offset, SEEK_SET) != 0) return std::string("short file ") + paths[p];
        const size_t bytes = (size_t)counts[p] * LIS_REC_BYTES;
        if (!up.copy(f.get(), dst, bytes))
            return std::string(paths[p]) + " holds fewer records than its header states";
        dst += bytes;
    }
    HIPCHK(plan::stream_wait(s));
    return "";
}

}  // namespace plan
