// Treecode.h -- drop-in C++ host API of the MI355X treecode.
//
// Rebuilds the reference class template Treecode<real_t>
// (code/Treecode.h:123-223) on top of the C ABI in bonsai2_amd.h, for float
// (the reference's only instantiation, code/TreecodeInstances.h:3) and for
// double (Particle4<double>, code/Particle4.h:106-125, computed in FP64), so a
// caller written against the reference keeps its call sequence
// (code/treecode.cu:61-165) unchanged:
//
//   typedef Treecode<float> Tree;
//   Tree tree(eps, theta);
//   tree.alloc(n);
//   for (int i = 0; i < n; i++) {
//     typename Tree::Particle pos, vel;
//     pos.x() = ...; pos.mass() = ...;
//     tree.h_ptclPos[i] = pos; tree.h_ptclVel[i] = vel;
//   }
//   tree.ptcl_h2d();
//   tree.buildTree(nLeaf);
//   tree.computeMultipoles();
//   tree.makeGroups(5, nGroup);
//   const double4 interactions = tree.computeForces();
//   tree.moveParticles();
//   tree.computeEnergies();
//
// The vector types (double4, float4, int2, uint4) are HIP's, from
// <hip/hip_vector_types.h>; a host compiler gets them with the flags
// `hipconfig --cpp_config` prints.  Link libbonsai2_amd.so.
//
// Public members as in the reference: theta, eps2, grav_potential; the pinned
// host arrays h_ptclPos / h_ptclVel (host_mem, code/cudamem.h:64-96) and
// ptcl0; the device arrays d_ptclPos, d_ptclVel, d_ptclPos_tmp, d_ptclAcc,
// d_domain, d_level_begIdx, d_cellDataList, d_cellSize, d_cellMonopole,
// d_cellQuad0, d_cellQuad1, d_groupList as views that read the context's
// current address (operator T*, d(), size(), d2h()).  ptcl_d2h() copies
// d_ptclPos / d_ptclVel as they are: tree order after buildTree, as in the
// reference (code/Treecode.h:200-204); ptcl_d2h_input() returns input order.
//
// Differences, all additive: errors throw bonsai2::Error instead of exit()
// (code/cutil.h:772-780); moveParticles() is the first half of a
// kick-drift-kick step with the member dt, closed by the next force
// evaluation (the reference's is a stub, code/moveParticles.cu:3-7);
// computeEnergies() fills ekin / epot in FP64, evaluating the forces first
// when the particles moved (a stub in the reference,
// code/computeEnergies.cu:3-7).
#pragma once

#include <hip/hip_vector_types.h>

#include <cmath>
#include <cstddef>
#include <new>
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

// Pinned host array, host_mem<T> of code/cudamem.h:64-96.
template <typename T>
struct host_mem {
  size_t n = 0;
  T *ptr = nullptr;
  host_mem() {}
  ~host_mem() { free(); }
  host_mem(const host_mem &) = delete;
  host_mem &operator=(const host_mem &) = delete;
  void alloc(size_t size) {
    free();
    void *p = nullptr;
    if (bt_host_alloc(&p, size * sizeof(T)) != 0) throw Error(BT_ERR_HIP, "bt_host_alloc failed");
    ptr = static_cast<T *>(p);
    n = size;
    for (size_t i = 0; i < n; i++) new (ptr + i) T();
  }
  void free() {
    if (ptr) bt_host_free(ptr);
    ptr = nullptr;
    n = 0;
  }
  void realloc(size_t size) {
    if (size > n) alloc(size);
  }
  operator T *() { return ptr; }
  const T &operator[](size_t i) const { return ptr[i]; }
  T &operator[](size_t i) { return ptr[i]; }
};

// A device array of the context (cuda_mem<T>, code/cudamem.h:5-61), read
// through the C ABI at each use: buildTree moves the particle arrays.
template <typename T>
class device_view {
  bt_ctx *const *ctx_;
  int which_;

  void *query(int64_t *count) const {
    void *p = nullptr;
    if (bt_get_device_buffer(*ctx_, which_, &p, count) != 0)
      throw Error(BT_ERR_STATE, std::string("device buffer: ") + bt_last_error(*ctx_));
    return p;
  }

 public:
  device_view(bt_ctx *const *ctx, int which) : ctx_(ctx), which_(which) {}
  T *d() const { return static_cast<T *>(query(nullptr)); }
  operator T *() const { return d(); }
  size_t size() const {
    int64_t k = 0;
    query(&k);
    return (size_t)k;
  }
  void d2h(T *host, int count = -1) const {
    const int64_t k = count < 0 ? (int64_t)size() : (int64_t)count;
    const int rc = bt_copy_device_buffer(*ctx_, which_, host, k);
    if (rc != 0) throw Error(rc, std::string("d2h: ") + bt_last_error(*ctx_));
  }
};

}  // namespace bonsai2

namespace bonsai2 {

// vec<N, T>::type as the reference spells it (code/Particle4.h:8-19)
template <int N, typename T> struct vec;
template <> struct vec<4, float> { typedef float4 type; };
template <> struct vec<4, double> { typedef double4 type; };
template <> struct vec<2, float> { typedef float2 type; };
template <> struct vec<2, double> { typedef double2 type; };

// The C ABI per real type: Treecode<float> and Treecode<double>.
namespace detail {
inline int create(bt_ctx **c, float eps, float theta, int ncrit) {
  return bt_create(c, eps, theta, ncrit);
}
inline int create(bt_ctx **c, double eps, double theta, int ncrit) {
  return bt_create_f64(c, eps, theta, ncrit);
}
inline int set_params(bt_ctx *c, float eps, float theta) { return bt_set_params(c, eps, theta); }
inline int set_params(bt_ctx *c, double eps, double theta) {
  return bt_set_params_f64(c, eps, theta);
}
inline int set_particles(bt_ctx *c, const float *p, const float *v) {
  return bt_set_particles(c, p, v);
}
inline int set_particles(bt_ctx *c, const double *p, const double *v) {
  return bt_set_particles_f64(c, p, v);
}
inline int get_particles(bt_ctx *c, float *p, float *v, int order) {
  return bt_get_particles(c, p, v, order);
}
inline int get_particles(bt_ctx *c, double *p, double *v, int order) {
  return bt_get_particles_f64(c, p, v, order);
}
inline int get_acc(bt_ctx *c, float *a, int order) { return bt_get_acc(c, a, order); }
inline int get_acc(bt_ctx *c, double *a, int order) { return bt_get_acc_f64(c, a, order); }
inline int move(bt_ctx *c, float dt) { return bt_move_particles(c, dt); }
inline int move(bt_ctx *c, double dt) { return bt_move_particles_f64(c, dt); }
}  // namespace detail

}  // namespace bonsai2

// Treecode<float> and Treecode<double> (code/Treecode.h:123-223).
template <typename real_t>
struct Treecode {
  static_assert(sizeof(real_t) == 4 || sizeof(real_t) == 8, "Treecode<float> or Treecode<double>");
  typedef bonsai2::Particle4<real_t> Particle;
  typedef typename bonsai2::vec<4, real_t>::type real4_t;
  typedef typename bonsai2::vec<2, real_t>::type real2_t;

  real_t theta, eps2;
  real_t dt = real_t(1) / 64;        // step of the no-argument moveParticles()
  double grav_potential = 0.0;       // code/Treecode.h:162
  double ekin = 0.0, epot = 0.0;     // set by computeEnergies()

 private:
  bt_ctx *ctx_ = nullptr;
  int nPtcl = 0, nCrit = 64, nLeaf = 16, levelSplit = 1;

  void check(int rc, const char *what) const {
    if (rc != 0) throw bonsai2::Error(rc, std::string(what) + ": " + bt_last_error(ctx_));
  }
  void sync_params() {
    check(bonsai2::detail::set_params(ctx_, eps2 > real_t(0) ? std::sqrt(eps2) : real_t(0), theta),
          "set_params");
  }

 public:
  // code/Treecode.h:142-160
  bonsai2::host_mem<Particle> h_ptclPos, h_ptclVel;
  std::vector<Particle> ptcl0;
  bonsai2::device_view<Particle> d_ptclPos{&ctx_, BT_BUF_PTCL_POS};
  bonsai2::device_view<Particle> d_ptclVel{&ctx_, BT_BUF_PTCL_VEL};
  bonsai2::device_view<Particle> d_ptclPos_tmp{&ctx_, BT_BUF_PTCL_POS_TMP};
  bonsai2::device_view<Particle> d_ptclAcc{&ctx_, BT_BUF_PTCL_ACC};
  bonsai2::device_view<real4_t> d_domain{&ctx_, BT_BUF_DOMAIN};
  bonsai2::device_view<int2> d_level_begIdx{&ctx_, BT_BUF_LEVEL_BEGIDX};
  bonsai2::device_view<uint4> d_cellDataList{&ctx_, BT_BUF_CELL_DATA};
  bonsai2::device_view<real4_t> d_cellSize{&ctx_, BT_BUF_CELL_SIZE};
  bonsai2::device_view<real4_t> d_cellMonopole{&ctx_, BT_BUF_CELL_MONOPOLE};
  bonsai2::device_view<real4_t> d_cellQuad0{&ctx_, BT_BUF_CELL_QUAD0};
  bonsai2::device_view<real2_t> d_cellQuad1{&ctx_, BT_BUF_CELL_QUAD1};
  bonsai2::device_view<int2> d_groupList{&ctx_, BT_BUF_GROUP_LIST};

  int get_nPtcl() const { return nPtcl; }
  int get_nCrit() const { return nCrit; }
  int get_nLeaf() const { return nLeaf; }

  // code/Treecode.h:164-173
  Treecode(const real_t _eps = real_t(0.01), const real_t _theta = real_t(0.75),
           const int _ncrit = 64)
      : theta(_theta), eps2(_eps * _eps), nCrit(_ncrit) {
    const int rc = bonsai2::detail::create(&ctx_, _eps, _theta, _ncrit);
    if (rc != 0) throw bonsai2::Error(rc, "bt_create failed (no HIP device?)");
  }
  ~Treecode() { bt_destroy(ctx_); }
  Treecode(const Treecode &) = delete;
  Treecode &operator=(const Treecode &) = delete;

  // code/Treecode.h:175-198
  void alloc(const int n) {
    nPtcl = n;
    h_ptclPos.alloc(n);
    h_ptclVel.alloc(n);
    check(bt_alloc(ctx_, n, 1, nullptr), "alloc");
  }

  // code/Treecode.h:207-214 (also keeps the initial positions in ptcl0)
  void ptcl_h2d() {
    sync_params();
    check(bonsai2::detail::set_particles(ctx_, &h_ptclPos[0].data[0], &h_ptclVel[0].data[0]),
          "ptcl_h2d");
    ptcl0.assign(h_ptclPos.ptr, h_ptclPos.ptr + nPtcl);
  }

  // code/Treecode.h:200-204: d_ptclPos / d_ptclVel as they are (tree order
  // after buildTree, input order before the first build)
  void ptcl_d2h() {
    check(bonsai2::detail::get_particles(ctx_, &h_ptclPos[0].data[0], &h_ptclVel[0].data[0],
                                         BT_ORDER_TREE),
          "ptcl_d2h");
  }
  // the same particles in the order ptcl_h2d gave them
  void ptcl_d2h_input() {
    check(bonsai2::detail::get_particles(ctx_, &h_ptclPos[0].data[0], &h_ptclVel[0].data[0],
                                         BT_ORDER_INPUT),
          "ptcl_d2h_input");
  }

  void buildTree(const int _nLeaf = 16) {  // code/buildTree.cu:845
    nLeaf = _nLeaf;
    sync_params();   // theta and eps2 are public members the caller may change
    check(bt_build_tree(ctx_, _nLeaf), "buildTree");
  }
  void computeMultipoles() {  // code/computeMultipoles.cu:194
    sync_params();
    check(bt_compute_multipoles(ctx_), "computeMultipoles");
  }
  void makeGroups(int _levelSplit = 1, const int _nCrit = 64) {  // code/makeGroups.cu:236
    nCrit = _nCrit;
    levelSplit = _levelSplit;
    check(bt_make_groups(ctx_, _levelSplit, _nCrit), "makeGroups");
  }
  // code/computeForces.cu:622-727: {direct <n>, direct max, approx <n>, approx max}
  double4 computeForces(const bool INTCOUNT = true) {
    double st[4] = {0, 0, 0, 0};
    check(bt_compute_forces(ctx_, INTCOUNT ? 1 : 0, st), "computeForces");
    check(bt_get_potential(ctx_, &grav_potential), "grav_potential");
    return make_double4(st[0], st[1], st[2], st[3]);
  }
  // first half of a kick-drift-kick step with dt; the next force evaluation
  // applies the second half kick
  void moveParticles() { check(bonsai2::detail::move(ctx_, dt), "moveParticles"); }
  void computeEnergies() {
    check(bt_compute_energies(ctx_, &ekin, &epot), "computeEnergies");
    check(bt_get_potential(ctx_, &grav_potential), "grav_potential");
  }

  // The C-ABI context, for the calls of bonsai2_amd.h that extend the
  // reference (learned opening angle, leaf groups, captured step, autotuner).
  bt_ctx *handle() { return ctx_; }
  bt_ctx *context() { return ctx_; }

  // (ax, ay, az, phi) per particle (d_ptclAcc), in input order by default
  std::vector<Particle> get_acc(int order = BT_ORDER_INPUT) {
    std::vector<Particle> a(nPtcl);
    check(bonsai2::detail::get_acc(ctx_, &a[0].data[0], order), "get_acc");
    return a;
  }
};
