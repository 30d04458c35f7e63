// host_io.cpp -- initial conditions and tipsy I/O (host only, no device).
//
// Samplers use a counter-based generator (Philox4x32-10, Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC11): draw k of
// particle i depends only on (seed, i, k), so any subset of particles can be
// regenerated and the result does not depend on thread count or order.
//
// Plummer: the Aarseth-Henon-Wielen algorithm of the reference sampler
// (code/plummer.h:37-117): radius from the inverse cumulative mass
// r = (X^(-2/3) - 1)^(-1/2) (rejected when r >= rmax), isotropic direction,
// speed by von Neumann rejection on q^2 (1-q^2)^(7/2), Henon scaling 3pi/16,
// m = 1/n, centre of mass removed in FP64.
//
// Tipsy: header {double time; int nbodies, ndim, nsph, ndark, nstar} padded
// to 32 bytes, then records, native endian (code/tipsydefs.h:22-51,
// code/read_tipsy.h:84-168).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/bonsai2_amd.h"

namespace {

// Runs f(i) for i in [b, e) on the host's cores (contiguous chunks).
template <typename F>
void parallel_for(int b, int e, F f) {
  const int n = e - b;
  if (n <= 0) return;
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 64) nt = 64;
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  const int chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int lo = b + t * chunk, hi = std::min(e, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=]() {
      for (int i = lo; i < hi; i++) f(i);
    });
  }
  for (auto &x : th) x.join();
}

struct Philox {
  uint64_t seed;
  // 4 x 32 random bits for counter (a, b, c)
  void block(uint32_t a, uint32_t b, uint32_t c, uint32_t out[4]) const {
    uint32_t x0 = a, x1 = b, x2 = c, x3 = 0x5eed5eedu;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; r++) {
      const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
      const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
      const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
      const uint32_t y1 = (uint32_t)p1;
      const uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
      const uint32_t y3 = (uint32_t)p0;
      x0 = y0; x1 = y1; x2 = y2; x3 = y3;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
  }
};

// Stream of uniform doubles in [0, 1) for one (particle, purpose) pair.
struct Stream {
  const Philox *ph;
  uint32_t i, purpose, k = 0;
  uint32_t buf[4];
  int avail = 0;
  double next() {
    if (avail == 0) {
      ph->block(i, k++, purpose, buf);
      avail = 2;
    }
    const int o = (2 - avail) * 2;
    avail--;
    const uint64_t a = buf[o] >> 5, b = buf[o + 1] >> 6;
    return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
  }
};

void plummer_one(Stream &s, double rmax, double p[3], double v[3]) {
  double R;
  for (;;) {
    const double X1 = s.next();
    if (X1 <= 0.0) continue;
    R = 1.0 / std::sqrt(std::pow(X1, -2.0 / 3.0) - 1.0);
    if (R < rmax) break;
  }
  const double X2 = s.next(), X3 = s.next();
  const double Z = (1.0 - 2.0 * X2) * R;
  const double Rxy = std::sqrt(std::max(0.0, R * R - Z * Z));
  p[0] = Rxy * std::cos(2.0 * M_PI * X3);
  p[1] = Rxy * std::sin(2.0 * M_PI * X3);
  p[2] = Z;
  const double Ve = std::sqrt(2.0) * std::pow(1.0 + R * R, -0.25);
  double X4 = 0.0, X5 = 0.0;
  while (0.1 * X5 >= X4 * X4 * std::pow(1.0 - X4 * X4, 3.5)) {
    X4 = s.next();
    X5 = s.next();
  }
  const double V = Ve * X4;
  const double X6 = s.next(), X7 = s.next();
  const double Vz = (1.0 - 2.0 * X6) * V;
  const double Vxy = std::sqrt(std::max(0.0, V * V - Vz * Vz));
  v[0] = Vxy * std::cos(2.0 * M_PI * X7);
  v[1] = Vxy * std::sin(2.0 * M_PI * X7);
  v[2] = Vz;
}

void remove_com(int n, const std::vector<double> &m, std::vector<double> &x,
                std::vector<double> &v) {
  double mt = 0, xc[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
  for (int i = 0; i < n; i++) {
    mt += m[i];
    for (int d = 0; d < 3; d++) {
      xc[d] += m[i] * x[3 * i + d];
      vc[d] += m[i] * v[3 * i + d];
    }
  }
  for (int d = 0; d < 3; d++) {
    xc[d] /= mt;
    vc[d] /= mt;
  }
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      x[3 * i + d] -= xc[d];
      v[3 * i + d] -= vc[d];
    }
}

void pack(int n, const std::vector<double> &m, const std::vector<double> &x,
          const std::vector<double> &v, float *pos4, float *vel4) {
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      pos4[4 * i + d] = (float)x[3 * i + d];
      vel4[4 * i + d] = (float)v[3 * i + d];
    }
    pos4[4 * i + 3] = (float)m[i];
    vel4[4 * i + 3] = (float)m[i];
  }
}

struct TipsyHeader {
  double time;
  int nbodies, ndim, nsph, ndark, nstar;
  int pad;
};
struct TipsyDark {
  float mass, pos[3], vel[3], eps;
  int phi;
};
struct TipsyStar {
  float mass, pos[3], vel[3], metals, tform, eps;
  int phi;
};
struct TipsyGas {
  float mass, pos[3], vel[3], rho, temp, hsmooth, metals, phi;
};
static_assert(sizeof(TipsyHeader) == 32, "tipsy header is 32 bytes");
static_assert(sizeof(TipsyDark) == 36, "tipsy dark record is 36 bytes");
static_assert(sizeof(TipsyStar) == 44, "tipsy star record is 44 bytes");
static_assert(sizeof(TipsyGas) == 48, "tipsy gas record is 48 bytes");

}  // namespace

extern "C" {

int bt_ic_plummer(int n, uint64_t seed, double rmax, float *pos4, float *vel4) {
  if (n <= 0 || !pos4 || !vel4 || !(rmax > 0.0)) return BT_ERR_ARG;
  const Philox ph{seed};
  std::vector<double> m(n, 1.0 / n), x(3 * (size_t)n), v(3 * (size_t)n);
  const double conv = 3.0 * M_PI / 16.0;
  parallel_for(0, n, [&](int i) {
    Stream s{&ph, (uint32_t)i, 1u};
    double p[3], q[3];
    plummer_one(s, rmax, p, q);
    for (int d = 0; d < 3; d++) {
      x[3 * (size_t)i + d] = p[d] * conv;
      v[3 * (size_t)i + d] = q[d] / std::sqrt(conv);
    }
  });
  remove_com(n, m, x, v);
  pack(n, m, x, v, pos4, vel4);
  return 0;
}

int bt_ic_clustered(int n, int nclust, uint64_t seed, double box, double mmin, double a0,
                    float *pos4, float *vel4) {
  if (n <= 0 || nclust <= 0 || nclust > n || !pos4 || !vel4 || !(mmin > 0.0) || !(a0 > 0.0))
    return BT_ERR_ARG;
  const Philox ph{seed};
  std::vector<double> cm(nclust), cx(3 * (size_t)nclust);
  double mt = 0.0;
  for (int k = 0; k < nclust; k++) {
    Stream s{&ph, (uint32_t)k, 7u};
    cm[k] = std::exp(std::log(mmin) * (1.0 - s.next()));  // log-uniform in [mmin, 1]
    for (int d = 0; d < 3; d++) cx[3 * k + d] = box * (2.0 * s.next() - 1.0);
    mt += cm[k];
  }
  for (int k = 0; k < nclust; k++) cm[k] /= mt;
  // particle counts proportional to mass (largest remainder), at least 1 each
  std::vector<int> cnt(nclust);
  int used = 0;
  for (int k = 0; k < nclust; k++) {
    cnt[k] = std::max(1, (int)std::floor(cm[k] * n));
    used += cnt[k];
  }
  for (int k = 0; used < n; k = (k + 1) % nclust) { cnt[k]++; used++; }
  for (int k = nclust - 1; used > n; k = (k + nclust - 1) % nclust)
    if (cnt[k] > 1) { cnt[k]--; used--; }
  std::vector<int> first(nclust + 1, 0);
  for (int k = 0; k < nclust; k++) first[k + 1] = first[k] + cnt[k];
  std::vector<double> m(n, 1.0 / n), x(3 * (size_t)n), v(3 * (size_t)n);
  const double conv = 3.0 * M_PI / 16.0;
  std::vector<int> owner(n);
  for (int k = 0; k < nclust; k++)
    for (int i = first[k]; i < first[k + 1]; i++) owner[i] = k;
  parallel_for(0, n, [&](int i) {
    const int k = owner[i];
    const double mk = (double)cnt[k] / n;          // cluster mass carried by its particles
    const double ak = a0 * std::cbrt(cm[k]);       // Henon length of the cluster
    const double vs = std::sqrt(mk / ak);          // velocity scale sqrt(G M / a)
    Stream s{&ph, (uint32_t)i, 3u};
    double p[3], q[3];
    plummer_one(s, 22.8, p, q);
    for (int d = 0; d < 3; d++) {
      x[3 * (size_t)i + d] = cx[3 * k + d] + p[d] * conv * ak;
      v[3 * (size_t)i + d] = q[d] / std::sqrt(conv) * vs;
    }
  });
  remove_com(n, m, x, v);
  pack(n, m, x, v, pos4, vel4);
  return 0;
}

int bt_ic_halo_disk(int n, uint64_t seed, double mh, double rd, double z0, float *pos4,
                    float *vel4) {
  if (n <= 1 || !pos4 || !vel4 || !(mh > 0.0 && mh < 1.0) || !(rd > 0.0) || !(z0 > 0.0))
    return BT_ERR_ARG;
  const Philox ph{seed};
  const int nh = (int)std::lround(mh * n);
  std::vector<double> m(n, 1.0 / n), x(3 * (size_t)n), v(3 * (size_t)n);
  const double conv = 3.0 * M_PI / 16.0;
  const double ah = conv;              // halo: Henon-unit Plummer of mass mh
  const double md = 1.0 - mh;
  parallel_for(0, nh, [&](int i) {
    Stream s{&ph, (uint32_t)i, 11u};
    double p[3], q[3];
    plummer_one(s, 22.8, p, q);
    for (int d = 0; d < 3; d++) {
      x[3 * (size_t)i + d] = p[d] * conv;
      v[3 * (size_t)i + d] = q[d] / std::sqrt(conv) * std::sqrt(mh);
    }
  });
  parallel_for(nh, n, [&](int i) {
    Stream s{&ph, (uint32_t)i, 13u};
    double R;
    do {
      // surface density ~ exp(-R/rd): R is Gamma(2, rd)
      const double u1 = 1.0 - s.next(), u2 = 1.0 - s.next();
      R = -rd * std::log(u1 * u2);
    } while (R > 10.0 * rd);
    const double phi = 2.0 * M_PI * s.next();
    const double z = z0 * std::atanh(std::max(-0.999999, std::min(0.999999, 2.0 * s.next() - 1.0)));
    x[3 * (size_t)i] = R * std::cos(phi);
    x[3 * (size_t)i + 1] = R * std::sin(phi);
    x[3 * (size_t)i + 2] = z;
    // circular speed from the enclosed halo and disk mass (spherical estimate)
    const double r2 = R * R;
    const double mhalo = mh * std::pow(r2 / (r2 + ah * ah), 1.5);
    const double mdisk = md * (1.0 - (1.0 + R / rd) * std::exp(-R / rd));
    const double vc = R > 0.0 ? std::sqrt((mhalo + mdisk) / R) : 0.0;
    v[3 * (size_t)i] = -vc * std::sin(phi);
    v[3 * (size_t)i + 1] = vc * std::cos(phi);
    v[3 * (size_t)i + 2] = 0.0;
  });
  remove_com(n, m, x, v);
  pack(n, m, x, v, pos4, vel4);
  return 0;
}

int bt_read_tipsy(const char *path, int *n, float *pos4, float *vel4, int cap) {
  if (!path || !n) return BT_ERR_ARG;
  FILE *f = std::fopen(path, "rb");
  if (!f) return BT_ERR_IO;
  TipsyHeader h;
  if (std::fread(&h, sizeof(h), 1, f) != 1 || h.nbodies < 0 || h.nsph < 0 || h.ndark < 0 ||
      h.nstar < 0 || h.nsph + h.ndark + h.nstar != h.nbodies) {
    std::fclose(f);
    return BT_ERR_IO;
  }
  *n = h.nbodies;
  if (!pos4) {
    std::fclose(f);
    return 0;
  }
  if (!vel4 || cap < h.nbodies) {
    std::fclose(f);
    return BT_ERR_ARG;
  }
  int k = 0;
  for (int i = 0; i < h.nsph; i++, k++) {
    TipsyGas g;
    if (std::fread(&g, sizeof(g), 1, f) != 1) { std::fclose(f); return BT_ERR_IO; }
    for (int d = 0; d < 3; d++) { pos4[4 * k + d] = g.pos[d]; vel4[4 * k + d] = g.vel[d]; }
    pos4[4 * k + 3] = vel4[4 * k + 3] = g.mass;
  }
  for (int i = 0; i < h.ndark; i++, k++) {
    TipsyDark d0;
    if (std::fread(&d0, sizeof(d0), 1, f) != 1) { std::fclose(f); return BT_ERR_IO; }
    for (int d = 0; d < 3; d++) { pos4[4 * k + d] = d0.pos[d]; vel4[4 * k + d] = d0.vel[d]; }
    pos4[4 * k + 3] = vel4[4 * k + 3] = d0.mass;
  }
  for (int i = 0; i < h.nstar; i++, k++) {
    TipsyStar s;
    if (std::fread(&s, sizeof(s), 1, f) != 1) { std::fclose(f); return BT_ERR_IO; }
    for (int d = 0; d < 3; d++) { pos4[4 * k + d] = s.pos[d]; vel4[4 * k + d] = s.vel[d]; }
    pos4[4 * k + 3] = vel4[4 * k + 3] = s.mass;
  }
  std::fclose(f);
  return 0;
}

int bt_write_tipsy(const char *path, int n, const float *pos4, const float *vel4, double time,
                   float eps) {
  if (!path || n < 0 || (n > 0 && (!pos4 || !vel4))) return BT_ERR_ARG;
  FILE *f = std::fopen(path, "wb");
  if (!f) return BT_ERR_IO;
  TipsyHeader h;
  std::memset(&h, 0, sizeof(h));
  h.time = time;
  h.nbodies = n;
  h.ndim = 3;
  h.ndark = n;
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  for (int i = 0; ok && i < n; i++) {
    TipsyDark d;
    d.mass = pos4[4 * i + 3];
    for (int k = 0; k < 3; k++) { d.pos[k] = pos4[4 * i + k]; d.vel[k] = vel4[4 * i + k]; }
    d.eps = eps;
    d.phi = i;
    ok = std::fwrite(&d, sizeof(d), 1, f) == 1;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : BT_ERR_IO;
}

}  // extern "C"
