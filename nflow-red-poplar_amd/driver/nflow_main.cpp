// nflow_main.cpp -- command-line driver over the C ABI, following the
// reference's program flow (main.cpp:50-351): parameters, decomposition,
// initialise, run maxitr steps, read arF back and print one slice of it; plus
// the CUDA path's outputs (kernel.cu): the residual log iteration.dat
// (150-151, 232-248) and result.csv / section.csv (972-1016).
//
// Without options it runs the reference's shipped case: a 10^3 cavity, MRT +
// LES (Cs = 0), Re = 1000, lid 0.1, 199 steps, 3x3x3 tiles.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nflow.h"

static void usage() {
  std::printf(
      "usage: nflow_main [options]\n"
      "  --n N | --nx N --ny N --nz N   lattice size (default 10^3)\n"
      "  --steps N                      time steps after initialisation (default 199)\n"
      "  --re R --lid U                 Reynolds number and lid velocity; tau = 3 U h / Re + 0.5\n"
      "  --tau T                        relaxation time (overrides --re)\n"
      "  --srt | --mrt                  collision model (default MRT)\n"
      "  --les CS                       Smagorinsky constant (default 0, as the reference)\n"
      "  --no-les                       disable the LES pass\n"
      "  --storage f32|f16|bf16         population storage (default f32)\n"
      "  --channel                      square-duct channel instead of the cavity\n"
      "  --pipe [R CY CZ]               circular pipe (solid outside radius R about (CY, CZ); default 20 25 25)\n"
      "  --decomp X Y Z                 sub-block counts (default 3 3 3, common.h:17-19)\n"
      "  --init-rest                    start every node at u = 0 (default: lid nodes at U)\n"
      "  --residual-every K             print and log the residual every K steps (iteration.dat)\n"
      "  --csv                          write result.csv and section.csv at the end\n"
      "  --vtu FILE                     write the final fields as a VTK unstructured grid\n"
      "  --restart FILE                 continue from a checkpoint file (same extents)\n"
      "  --checkpoint FILE              save a checkpoint file at the end\n"
      "  --no-print                     skip the arF slice print\n"
      "  --block B                      threads per workgroup of the one-step launches (64..512)\n"
      "  --temporal H                   two steps per launch in y chunks of H planes (FP32 cavity, channel);\n"
      "                                 0 = one step per launch, -1 = the library's choice (default)\n");
}

static int check(int rc, const char* what) {
  if (rc < 0) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, nflow_last_error());
    std::exit(1);
  }
  return rc;
}

int main(int argc, char** argv) {
  nflow_input in;
  nflow_input_defaults(&in);
  int xd = 3, yd = 3, zd = 3, residual_every = 0, block = 0, temporal = -1;
  bool csv = false, print = true, tau_set = false;
  std::string vtu, restart, checkpoint;
  double re = 1000.0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--n") in.nx = in.ny = in.nz = in.h = std::atoi(next());
    else if (a == "--nx") in.nx = std::atoi(next());
    else if (a == "--ny") in.ny = std::atoi(next());
    else if (a == "--nz") in.nz = std::atoi(next());
    else if (a == "--steps") in.maxitr = std::atoi(next());
    else if (a == "--re") re = std::atof(next());
    else if (a == "--lid") in.ref_vel[0] = std::atof(next());
    else if (a == "--tau") { in.tau = std::atof(next()); tau_set = true; }
    else if (a == "--srt") in.enable_mrt = 0;
    else if (a == "--mrt") in.enable_mrt = 1;
    else if (a == "--les") { in.smagorinsky = std::atof(next()); in.enable_les = 1; }
    else if (a == "--no-les") in.enable_les = 0;
    else if (a == "--storage") {
      std::string s = next();
      in.storage = s == "f16" ? NFLOW_STORAGE_F16 : s == "bf16" ? NFLOW_STORAGE_BF16 : NFLOW_STORAGE_F32;
    } else if (a == "--channel") in.casetype = NFLOW_CHANNEL;
    else if (a == "--pipe") {
      in.casetype = NFLOW_PIPE;
      if (i + 3 < argc && argv[i + 1][0] != '-') {
        in.pipe_radius = std::atof(next());
        in.pipe_cy = std::atof(next());
        in.pipe_cz = std::atof(next());
      }
    }
    else if (a == "--decomp") { xd = std::atoi(next()); yd = std::atoi(next()); zd = std::atoi(next()); }
    else if (a == "--init-rest") in.init_lid_velocity = 0;
    else if (a == "--residual-every") residual_every = std::atoi(next());
    else if (a == "--csv") csv = true;
    else if (a == "--vtu") vtu = next();
    else if (a == "--restart") restart = next();
    else if (a == "--checkpoint") checkpoint = next();
    else if (a == "--no-print") print = false;
    else if (a == "--block") block = std::atoi(next());
    else if (a == "--temporal") temporal = std::atoi(next());
    else { usage(); return a == "--help" ? 0 : 2; }
  }
  if (in.h != in.nx) in.h = in.nx;
  if (!tau_set) in.tau = 3.0 * in.ref_vel[0] * in.h / re + 0.5;  // main.cpp:81

  std::printf(" - %dx%dx%d %s, %s%s, tau = %.6f, Cs = %.3f, storage = %d\n", in.nx, in.ny, in.nz,
              in.casetype == NFLOW_CAVITY ? "cavity" : in.casetype == NFLOW_PIPE ? "pipe" : "channel", in.enable_mrt ? "MRT" : "SRT",
              in.enable_les ? "+LES" : "", in.tau, in.smagorinsky, in.storage);

  std::vector<int> slice(6 * (size_t)xd * yd * zd);
  auto* sl = reinterpret_cast<int(*)[3][2]>(slice.data());
  const int ntile = check(nflow_decomposition(&in, xd, yd, zd, sl, xd * yd * zd), "decomposition");
  std::printf("Number of Tiles participating = %d\n", ntile);

  nflow_solver* s = nullptr;
  check(nflow_create(&in, ntile, sl, &s), "nflow_create");
  if (block > 0 || temporal != -1) {
    nflow_tuning t = {block > 0 ? block : (in.storage == NFLOW_STORAGE_F32 ? 256 : 512), 0, 0, 10, temporal, 0.0};
    check(nflow_set_tuning(s, &t), "nflow_set_tuning");
  }
  check(nflow_initialize(s), "nflow_initialize");
  if (!restart.empty()) {
    check(nflow_load_checkpoint(s, restart.c_str()), "nflow_load_checkpoint");
    std::printf(" - restarted from %s at step %lld\n", restart.c_str(), nflow_steps_done(s));
  }

  FILE* log = nullptr;
  if (residual_every > 0) {
    log = std::fopen("iteration.dat", "wt");
    if (log) std::fprintf(log, "%s\n", "iter, error");
  }
  auto t0 = std::chrono::steady_clock::now();
  double dev_ms = 0.0;
  int done = 0;
  while (done < in.maxitr) {
    const int k = residual_every > 0 ? std::min(residual_every, in.maxitr - done) : in.maxitr - done;
    check(nflow_run(s, k), "nflow_run");
    double ms = 0.0;
    check(nflow_timing(s, &ms, nullptr), "nflow_timing");
    dev_ms += ms * k;
    done += k;
    if (residual_every > 0) {
      double mass, mom[3], err;
      check(nflow_diagnostics(s, &mass, mom, &err), "nflow_diagnostics");
      std::printf("ITER=%d | ERROR=%.5e\n", done, err);
      if (log) std::fprintf(log, "%d %.5e\n", done, err);
    }
  }
  check(nflow_synchronize(s), "nflow_synchronize");
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (log) std::fclose(log);
  const double nodes = (double)in.nx * in.ny * in.nz;
  std::printf(" - Elapsed time(GPU) = %fs | h=%d | %d steps | %.1f MLUPS (device %.1f)\n", wall, in.h, in.maxitr,
              nodes * in.maxitr / wall / 1e6, dev_ms > 0 ? nodes * in.maxitr / (dev_ms * 1e-3) / 1e6 : 0.0);

  const int px = in.nx + 2, pz = in.nz + 2;
  std::vector<float> arF((size_t)(in.ny + 2) * pz * px * NFLOW_DIRECTION_SIZE);
  check(nflow_read_arF(s, arF.data()), "nflow_read_arF");
  if (print) {
    // main.cpp:321-348: direction 9, z-plane k = 2, rows y = ny-1 .. 0
    const int k = std::min(2, in.nz - 1);
    std::printf("---------------------------------\n");
    for (int j = in.ny - 1; j >= 0; --j) {
      for (int i = 0; i < in.nx; ++i) {
        const size_t id = (((size_t)(j + 1) * pz + (k + 1)) * px + (i + 1)) * NFLOW_DIRECTION_SIZE + 9;
        std::printf("%8.6f ", arF[id]);
      }
      std::printf("\n");
    }
    std::printf("\n---------------------------------\n");
  }
  if (!checkpoint.empty()) check(nflow_save_checkpoint(s, checkpoint.c_str()), "nflow_save_checkpoint");
  if (!vtu.empty()) {
    const size_t n = (size_t)nodes;
    std::vector<float> rho(n), ux(n), uy(n), uz(n);
    std::vector<int> type(n), dirs(n);
    check(nflow_read_macro(s, rho.data(), ux.data(), uy.data(), uz.data()), "nflow_read_macro");
    check(nflow_read_node_type(s, type.data(), dirs.data()), "nflow_read_node_type");
    check(nflow_write_vtu(vtu.c_str(), in.nx, in.ny, in.nz, rho.data(), ux.data(), uy.data(), uz.data(), type.data()),
          "nflow_write_vtu");
  }
  if (csv) {
    const size_t n = (size_t)nodes;
    std::vector<float> rho(n), ux(n), uy(n), uz(n);
    check(nflow_read_macro(s, rho.data(), ux.data(), uy.data(), uz.data()), "nflow_read_macro");
    FILE* f = std::fopen("result.csv", "wt");
    if (!f) { std::perror("result.csv"); return 1; }
    std::fprintf(f, "%s\n", "x,y,z,rho,u,v,w");
    for (size_t id = 0; id < n; ++id) {
      const int x = (int)(id % in.nx), z = (int)((id / in.nx) % in.nz), y = (int)(id / ((size_t)in.nx * in.nz));
      std::fprintf(f, "%d, %d, %d, %lf, %lf, %lf, %lf\n", x, y, z, rho[id], ux[id], uy[id], uz[id]);
    }
    std::fclose(f);
    // vertical centre line u/U (the reference hard-codes x = z = 49 and U = 0.1)
    FILE* g = std::fopen("section.csv", "wt");
    if (!g) { std::perror("section.csv"); return 1; }
    std::fprintf(g, "%s\n", "y/Y,u");
    const int xc = in.nx / 2, zc = in.nz / 2;
    for (int y = 0; y < in.ny; ++y) {
      const size_t id = (size_t)xc + (size_t)in.nx * ((size_t)zc + (size_t)in.nz * y);
      std::fprintf(g, "%d, %lf\n", y, in.ref_vel[0] != 0.0 ? ux[id] / in.ref_vel[0] : 0.0);
    }
    std::fclose(g);
  }
  nflow_destroy(s);
  return 0;
}
