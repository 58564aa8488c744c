// boundness.hip — placeholder, filled below
