// Treecode.h -- drop-in C++ host API of the MI355X treecode.
//
// Rebuilds the reference class template Treecode<real_t>
// (code/Treecode.h:123-223, instantiated for float only,
// code/TreecodeInstances.h:3) on top of the C ABI in bonsai2_amd.h, so a
// caller written against the reference keeps its call sequence
// (code/treecode.cu:61-165):
//
//   Treecode<float> tree(eps, theta);
//   tree.alloc(n);
//   ... fill tree.h_ptclPos[i] / tree.h_ptclVel[i] ...
//   tree.ptcl_h2d();
//   tree.buildTree(nLeaf);
//   tree.computeMultipoles();
//   tree.makeGroups(5, nGroup);
//   double4 interactions = tree.computeForces();
//   tree.moveParticles();
//   tree.computeEnergies();
//
// Differences, all additive: errors throw bonsai2::Error instead of exit()
// (code/cutil.h:772-780); moveParticles() integrates with the member dt
// (the reference's is a stub, code/moveParticles.cu:3-7); computeEnergies()
// fills ekin/epot (a stub in the reference, code/computeEnergies.cu:3-7);
// d_ptclAcc is read with get_acc().  Host only: include with any C++11
// compiler and link libbonsai2_amd.so.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "bonsai2_amd.h"

namespace bonsai2 {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Particle record of code/Particle4.h:59-104: (x, y, z, mass) / (vx, vy, vz, mass).
template <typename T>
struct Particle4 {
  T data[4];
  T &x() { return data[0]; }
  T &y() { return data[1]; }
  T &z() { return data[2]; }
  T &mass() { return data[3]; }
  T x() const { return data[0]; }
  T y() const { return data[1]; }
  T z() const { return data[2]; }
  T mass() const { return data[3]; }
};

struct double4_t {
  double x, y, z, w;
};

}  // namespace bonsai2

template <typename real_t>
struct Treecode;

template <>
struct Treecode<float> {
  typedef bonsai2::Particle4<float> Particle;

  float theta, eps2;
  float dt = 1.0f / 64;              // step for the no-argument moveParticles()
  double grav_potential = 0.0;       // code/Treecode.h:162
  double ekin = 0.0, epot = 0.0;     // set by computeEnergies()
  std::vector<Particle> h_ptclPos, h_ptclVel;  // host particle arrays (code/Treecode.h:142)

 private:
  bt_ctx *ctx_ = nullptr;
  int nPtcl = 0, nCrit = 64, nLeaf = 16;
  bool moved_ = false;

  void check(int rc, const char *what) const {
    if (rc != 0) throw bonsai2::Error(rc, std::string(what) + ": " + bt_last_error(ctx_));
  }

 public:
  int get_nPtcl() const { return nPtcl; }
  int get_nCrit() const { return nCrit; }
  int get_nLeaf() const { return nLeaf; }
  bt_ctx *context() { return ctx_; }

  // code/Treecode.h:164-173
  Treecode(const float _eps = 0.01f, const float _theta = 0.75f, const int _ncrit = 64)
      : theta(_theta), eps2(_eps * _eps), nCrit(_ncrit) {
    const int rc = bt_create(&ctx_, _eps, _theta, _ncrit);
    if (rc != 0) throw bonsai2::Error(rc, "bt_create failed (no HIP device?)");
  }
  ~Treecode() { bt_destroy(ctx_); }
  Treecode(const Treecode &) = delete;
  Treecode &operator=(const Treecode &) = delete;

  // code/Treecode.h:175-198
  void alloc(const int n) {
    nPtcl = n;
    h_ptclPos.assign(n, Particle());
    h_ptclVel.assign(n, Particle());
    check(bt_alloc(ctx_, n, 1, nullptr), "alloc");
  }

  // code/Treecode.h:207-214
  void ptcl_h2d() {
    float eps = 0.f;
    if (eps2 > 0.f) eps = std::sqrt(eps2);
    check(bt_set_params(ctx_, eps, theta), "set_params");
    check(bt_set_particles(ctx_, &h_ptclPos[0].data[0], &h_ptclVel[0].data[0]), "ptcl_h2d");
  }

  // code/Treecode.h:200-204 (particles come back in input order)
  void ptcl_d2h() {
    check(bt_get_particles(ctx_, &h_ptclPos[0].data[0], &h_ptclVel[0].data[0],
                           BT_ORDER_INPUT),
          "ptcl_d2h");
  }

  void buildTree(const int _nLeaf = 16) {  // code/buildTree.cu:845
    nLeaf = _nLeaf;
    if (moved_) ptcl_sync_params();
    check(bt_build_tree(ctx_, _nLeaf), "buildTree");
  }
  void computeMultipoles() { check(bt_compute_multipoles(ctx_), "computeMultipoles"); }
  void makeGroups(int levelSplit = 1, const int _nCrit = 64) {  // code/makeGroups.cu:236
    nCrit = _nCrit;
    check(bt_make_groups(ctx_, levelSplit, _nCrit), "makeGroups");
  }
  // code/computeForces.cu:622-727: {direct <n>, direct max, approx <n>, approx max}
  bonsai2::double4_t computeForces(const bool INTCOUNT = true) {
    double st[4] = {0, 0, 0, 0};
    check(bt_compute_forces(ctx_, INTCOUNT ? 1 : 0, st), "computeForces");
    check(bt_get_potential(ctx_, &grav_potential), "grav_potential");
    return bonsai2::double4_t{st[0], st[1], st[2], st[3]};
  }
  // half kick with the current forces, then drift (v += a dt/2, x += v dt)
  void moveParticles() {
    check(bt_move_particles(ctx_, dt), "moveParticles");
    moved_ = true;
  }
  void computeEnergies() { check(bt_compute_energies(ctx_, &ekin, &epot), "computeEnergies"); }

  // The C-ABI context, for the calls of bonsai2_amd.h that extend the
  // reference (learned opening angle, leaf groups, captured step, autotuner).
  bt_ctx *handle() { return ctx_; }

  // (ax, ay, az, phi) per particle (reference d_ptclAcc), in input order by default
  std::vector<bonsai2::Particle4<float>> get_acc(int order = BT_ORDER_INPUT) {
    std::vector<bonsai2::Particle4<float>> a(nPtcl);
    check(bt_get_acc(ctx_, &a[0].data[0], order), "get_acc");
    return a;
  }

 private:
  void ptcl_sync_params() {
    float eps = 0.f;
    if (eps2 > 0.f) eps = std::sqrt(eps2);
    check(bt_set_params(ctx_, eps, theta), "set_params");
  }
};
