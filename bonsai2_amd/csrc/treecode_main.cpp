// treecode_main.cpp -- command-line driver, the counterpart of the reference's
// code/treecode.cu:3-174, written against the drop-in include/Treecode.h.
//
// Flags (code/treecode.cu:21-39): -i tipsy file, -n Plummer particle count,
// -s seed, -l nLeaf, -g nGroup (NCRIT), -o theta, -e eps; added: -t steps,
// -d dt, -w output tipsy snapshot.  Prints the reference's interaction
// statistics and TFLOP/s line (code/treecode.cu:149-158) plus FP64 energies.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/Treecode.h"

namespace {

void usage() {
  std::printf(
      "Usage\n"
      " -h  --help             Prints this help\n"
      " -i  --infile #         Input snapshot filename [tipsy format]\n"
      " -n  --plummer #        Generate plummer model with a given number of particles\n"
      " -s  --seed    #        Random seed [19810614]\n"
      " -l  --nleaf   #        Number of particles in leaf (16,24,32,48,64) [64]\n"
      " -g  --ngroup  #        Number of particles in group [64]\n"
      " -o  --theta   #        Opening angle [0.75]\n"
      " -e  --eps     #        Softening [0.05]\n"
      " -t  --steps   #        Leapfrog steps after the first force evaluation [0]\n"
      " -d  --dt      #        Time step [1/64]\n"
      " -w  --outfile #        Write the final state as a tipsy snapshot\n");
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

int main(int argc, char **argv) {
  std::string infile, outfile;
  long n = -1;
  unsigned long long seed = 19810614;
  int nleaf = 64, ngroup = 64, steps = 0;
  float theta = 0.75f, eps = 0.05f, dt = 1.0f / 64;
  if (argc < 2) {
    usage();
    return 0;
  }
  for (int k = 1; k < argc; k++) {
    const std::string a = argv[k];
    auto val = [&](const char *name) -> const char * {
      if (k + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", name);
        std::exit(2);
      }
      return argv[++k];
    };
    if (a == "-h" || a == "--help") { usage(); return 0; }
    else if (a == "-i" || a == "--infile") infile = val("-i");
    else if (a == "-n" || a == "--plummer") n = std::atol(val("-n"));
    else if (a == "-s" || a == "--seed") seed = std::strtoull(val("-s"), nullptr, 10);
    else if (a == "-l" || a == "--nleaf") nleaf = std::atoi(val("-l"));
    else if (a == "-g" || a == "--ngroup") ngroup = std::atoi(val("-g"));
    else if (a == "-o" || a == "--theta") theta = (float)std::atof(val("-o"));
    else if (a == "-e" || a == "--eps") eps = (float)std::atof(val("-e"));
    else if (a == "-t" || a == "--steps") steps = std::atoi(val("-t"));
    else if (a == "-d" || a == "--dt") dt = (float)std::atof(val("-d"));
    else if (a == "-w" || a == "--outfile") outfile = val("-w");
    else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
  }
  if (!(nleaf == 16 || nleaf == 24 || nleaf == 32 || nleaf == 48 || nleaf == 64)) {
    std::fprintf(stderr, "nleaf must be 16, 24, 32, 48 or 64\n");
    return 2;
  }
  std::vector<float> pos, vel;
  if (n > 0) {
    pos.resize(4 * n);
    vel.resize(4 * n);
    if (bt_ic_plummer((int)n, seed, 22.8, pos.data(), vel.data()) != 0) return 3;
    std::printf("Using Plummer model with nPtcl= %ld\n", n);
  } else if (!infile.empty()) {
    int ni = 0;
    if (bt_read_tipsy(infile.c_str(), &ni, nullptr, nullptr, 0) != 0) {
      std::fprintf(stderr, "cannot read %s\n", infile.c_str());
      return 3;
    }
    n = ni;
    pos.resize(4 * n);
    vel.resize(4 * n);
    if (bt_read_tipsy(infile.c_str(), &ni, pos.data(), vel.data(), ni) != 0) return 3;
  } else {
    usage();
    return 2;
  }
  try {
    Treecode<float> tree(eps, theta, ngroup);
    tree.dt = dt;
    tree.alloc((int)n);
    for (long i = 0; i < n; i++)
      for (int d = 0; d < 4; d++) {
        tree.h_ptclPos[i].data[d] = pos[4 * i + d];
        tree.h_ptclVel[i].data[d] = vel[4 * i + d];
      }
    tree.ptcl_h2d();
    const double t0 = now();
    tree.buildTree(nleaf);
    tree.computeMultipoles();
    tree.makeGroups(5, ngroup);
    const double t1 = now();
    const double4 st = tree.computeForces();
    const double dtw = now() - t1;
    std::fprintf(stderr,
                 " direct: <n>= %g  max= %g  approx: <n>=%g max= %g :: perf= %g TFLOP/s \n",
                 st.x, st.y, st.z, st.w, (st.x * 20 + st.z * 64) * n / dtw / 1e12);
    tree.computeEnergies();
    const double e0 = tree.ekin + tree.epot;
    std::fprintf(stderr, " grav potential= %g  E= %.10g (K= %g W= %g)\n", tree.grav_potential,
                 e0, tree.ekin, tree.epot);
    bt_ctx *c = tree.context();
    for (int s = 0; s < steps; s++) {
      if (bt_step(c, dt, nleaf, 5, ngroup) != 0) {
        std::fprintf(stderr, "step failed: %s\n", bt_last_error(c));
        return 4;
      }
    }
    if (steps > 0) {
      tree.computeEnergies();
      const double e1 = tree.ekin + tree.epot;
      std::fprintf(stderr, " after %d steps: E= %.10g  |dE/E|= %.3e\n", steps, e1,
                   std::fabs((e1 - e0) / e0));
    }
    std::fprintf(stderr, " steps are done in %g sec :: rate = %g MPtcl/sec\n", now() - t0,
                 n / 1e6 / (now() - t0));
    std::fprintf(stderr, " nLeaf= %d  nCrit= %d\n", tree.get_nLeaf(), tree.get_nCrit());
    if (!outfile.empty()) {
      tree.ptcl_d2h_input();
      if (bt_write_tipsy(outfile.c_str(), (int)n, &tree.h_ptclPos[0].data[0],
                         &tree.h_ptclVel[0].data[0], dt * steps, eps) != 0)
        return 5;
    }
  } catch (const bonsai2::Error &e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 10;
  }
  return 0;
}
