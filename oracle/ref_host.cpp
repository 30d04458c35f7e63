// ref_host.cpp -- thin C entry points around the reference's own host-only
// sources, compiled in place from the reference tree (never copied) into
// oracle/_ref/libref_host.so.  Test infrastructure only: it pins the
// oracle's Plummer restatement (code/plummer.h:9-137) and the product's
// tipsy writer against the reference reader (code/read_tipsy.h:84-168).
#include <cstring>
#include "plummer.h"
#include "read_tipsy.h"

extern "C" {

// Runs the reference Plummer constructor.  The cache file name points into a
// directory that does not exist, so plummer.h:15-34 finds no cache and
// :119-133 cannot write one.  n must be >= 64 (plummer.h:139 divides by n/64).
int ref_plummer(long n, unsigned seed, double *mass, double *pos, double *vel) {
  if (n < 64) return -1;
  const Plummer p(n, seed, "/nonexistent-bonsai-cache-dir/plummer.dat");
  for (long i = 0; i < n; i++) {
    mass[i] = p.mass[i];
    for (int d = 0; d < 3; d++) {
      pos[3 * i + d] = p.pos[i][d];
      vel[3 * i + d] = p.vel[i][d];
    }
  }
  return 0;
}

// Reads a tipsy file with the reference reader.  Returns NTotal; fills
// pos/vel (4 floats each, w = mass / eps as the reader stores them) when
// the buffers are non-null and hold at least cap particles.
int ref_read_tipsy(const char *fname, int cap, float *pos, float *vel, int *ids) {
  const ReadTipsy r{std::string(fname)};
  const int n = r.NTotal;
  if (pos && vel && n <= cap) {
    for (int i = 0; i < n; i++) {
      pos[4 * i] = r.positions[i].x; pos[4 * i + 1] = r.positions[i].y;
      pos[4 * i + 2] = r.positions[i].z; pos[4 * i + 3] = r.positions[i].w;
      vel[4 * i] = r.velocities[i].x; vel[4 * i + 1] = r.velocities[i].y;
      vel[4 * i + 2] = r.velocities[i].z; vel[4 * i + 3] = r.velocities[i].w;
      if (ids) ids[i] = r.IDs[i];
    }
  }
  return n;
}

}  // extern "C"
