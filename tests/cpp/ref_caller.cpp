// ref_caller.cpp -- a caller of the drop-in include/Treecode.h written to the
// call sequence of the reference driver, code/treecode.cu:61-165: the same
// Treecode<real_t> members in the same order and the reference's double4
// statistics.  The reference's Plummer class (code/plummer.h)
// is not part of the library, so the particles come from bt_ic_plummer.
//
// Usage: ref_caller N NLEAF NGROUP THETA EPS OUT [f64]
// With f64 the caller is Treecode<double> (particles converted from the
// float sample).  OUT (binary): double[8] = interactions.xyzw,
// grav_potential after computeForces, ekin, epot, grav_potential after
// computeEnergies; then real4[N] d_ptclAcc (PH order, read through the
// d_ptclAcc view after computeForces); then real4[N] h_ptclPos after
// moveParticles + computeEnergies + ptcl_d2h (tree order of the last build).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "Treecode.h"

template <typename real_t>
int run(int nPtcl, int nLeaf, int nGroup, real_t theta, real_t eps, const char *out) {
  std::vector<float> ppos(4 * (size_t)nPtcl), pvel(4 * (size_t)nPtcl);
  if (bt_ic_plummer(nPtcl, 19810614ull, 22.8, ppos.data(), pvel.data()) != 0) return 3;

  typedef Treecode<real_t> Tree;
  Tree tree(eps, theta);
  tree.alloc(nPtcl);
  for (int i = 0; i < nPtcl; i++) {
    typename Tree::Particle ptclPos, ptclVel;
    ptclPos.x() = ppos[4 * i + 0];
    ptclPos.y() = ppos[4 * i + 1];
    ptclPos.z() = ppos[4 * i + 2];
    ptclVel.x() = pvel[4 * i + 0];
    ptclVel.y() = pvel[4 * i + 1];
    ptclVel.z() = pvel[4 * i + 2];
    ptclVel.mass() = ppos[4 * i + 3];
    ptclPos.mass() = ppos[4 * i + 3];
    tree.h_ptclPos[i] = ptclPos;
    tree.h_ptclVel[i] = ptclVel;
  }
  {
    double mtot = 0.0;
    for (int i = 0; i < nPtcl; i++) mtot += tree.h_ptclPos[i].mass();
    std::fprintf(stderr, " Total mass = %g \n", mtot);
  }

  tree.ptcl_h2d();
  tree.buildTree(nLeaf);
  tree.computeMultipoles();
  tree.makeGroups(5, nGroup);
  const double4 interactions = tree.computeForces();
  const double pot_forces = tree.grav_potential;
  std::fprintf(stderr, " direct: <n>= %g  max= %g  approx: <n>=%g max= %g\n", interactions.x,
               interactions.y, interactions.z, interactions.w);
  std::vector<typename Tree::Particle> acc(nPtcl);
  tree.d_ptclAcc.d2h(acc.data());
  if (tree.d_ptclAcc.size() != (size_t)nPtcl || tree.d_cellDataList.size() == 0 ||
      tree.d_groupList.size() == 0 || tree.d_ptclAcc.d() == nullptr)
    return 4;
  tree.moveParticles();
  tree.computeEnergies();
  std::fprintf(stderr, " E= %.10g (K= %g W= %g)\n", tree.ekin + tree.epot, tree.ekin, tree.epot);
  tree.ptcl_d2h();

  FILE *f = std::fopen(out, "wb");
  if (!f) return 5;
  const double head[8] = {interactions.x, interactions.y, interactions.z, interactions.w,
                          pot_forces, tree.ekin, tree.epot, tree.grav_potential};
  std::fwrite(head, sizeof(double), 8, f);
  std::fwrite(acc.data(), sizeof(typename Tree::Particle), nPtcl, f);
  std::fwrite(&tree.h_ptclPos[0], sizeof(typename Tree::Particle), nPtcl, f);
  std::fclose(f);
  return 0;
}

int main(int argc, char *argv[]) {
  if (argc != 7 && argc != 8) {
    std::fprintf(stderr, "usage: %s N NLEAF NGROUP THETA EPS OUT [f64]\n", argv[0]);
    return 2;
  }
  const int nPtcl = std::atoi(argv[1]);
  const int nLeaf = std::atoi(argv[2]);
  const int nGroup = std::atoi(argv[3]);
  const double theta = std::atof(argv[4]);
  const double eps = std::atof(argv[5]);
  if (argc == 8 && std::strcmp(argv[7], "f64") == 0)
    return run<double>(nPtcl, nLeaf, nGroup, theta, eps, argv[6]);
  return run<float>(nPtcl, nLeaf, nGroup, (float)theta, (float)eps, argv[6]);
}
