// drive_step.cpp -- a C++ caller of the drag library's C-ABI, without Python: it builds a uniform
// periodic mesh, creates a context, sets up the seeded problem on the device and runs imex2+
// steps with the dk_stage records an AthenaK task body would pass (StageOf of the integration
// notes), then prints the FP64 diagnostic sums and a checksum of the exported particles.
//
//   drive_step <form 0|1> <steps> [graphs 0|1]
//
// Mesh: [0,1)^3 periodic, 16^3 cells in 2x2x2 blocks of 8^3, ng = 2.  Problem: ppc 2, one
// species t_s = 0.1, eps = 1, 1e-3 noise on gas and dust, seed 1; TSC, imex2+ with dc1, dt 1e-3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dustk.h"
#include "dustk_setup.h"

namespace {

// the parts of AthenaK's Driver and Mesh that a dust task reads
struct Driver {
  int nexp_stages = 3;
  double gam0[3], gam1[3], beta[3];
  double a_twid[3][3];
  double a_impl;
};
struct Mesh {
  double time = 0.0, dt = 1.0e-3;
};

// Driver::SetImEx2PlusCoefficients (src/driver/driver.cpp:278-300), gamma = 1 + 1/sqrt(2)
void SetImEx2Plus(Driver &d) {
  const double g = 1.707106781186547;
  d.gam0[0] = 1.0; d.gam1[0] = 0.0; d.beta[0] = g;
  d.gam0[1] = (2.0*g - 1.0)/(2.0*g*g);
  d.gam1[1] = 1.0 - d.gam0[1];
  d.beta[1] = 1.0/(2.0*g);
  d.gam0[2] = 1.0; d.gam1[2] = 0.0; d.beta[2] = 0.0;
  std::memset(d.a_twid, 0, sizeof(d.a_twid));
  d.a_twid[2][2] = (1.0 - 2.0*g*g)/2.0/g;
  d.a_impl = g;
}

// the stage record of a task body (as in the integration notes)
dk_stage StageOf(const Driver *d, const Mesh *pm, int stage) {
  dk_stage st{};
  st.stage = stage; st.nexp_stages = d->nexp_stages;
  st.time = pm->time; st.dt = pm->dt;
  for (int k = 0; k < 3; ++k) {st.gam0[k] = d->gam0[k]; st.gam1[k] = d->gam1[k];
                               st.beta[k] = d->beta[k];}
  st.a_twid22 = d->a_twid[2][2]; st.a_impl = d->a_impl;
  return st;
}

// LeftEdgeX (src/coordinates/cell_locations.hpp:24-27)
double LeftEdge(int ith, int n, double xmin, double xmax) {
  const double x = static_cast<double>(ith)/static_cast<double>(n);
  return (x*xmax - x*xmin) - (0.5*xmax - 0.5*xmin) + (0.5*xmin + 0.5*xmax);
}

int Fail(const char *what, int e) {
  std::fprintf(stderr, "drive_step: %s: %s\n", what, dk_strerror(e));
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: drive_step <form 0|1> <steps> [graphs 0|1]\n");
    return 2;
  }
  const int form = std::atoi(argv[1]), steps = std::atoi(argv[2]);
  const int graphs = (argc > 3) ? std::atoi(argv[3]) : 0;
  const int nb = 2, nx = 8, ng = 2;
  dk_mesh mesh{};
  mesh.nmb = nb*nb*nb; mesh.nx1 = mesh.nx2 = mesh.nx3 = nx; mesh.ng = ng;
  mesh.x1min = mesh.x2min = mesh.x3min = 0.0;
  mesh.x1max = mesh.x2max = mesh.x3max = 1.0;
  mesh.nmbx1 = mesh.nmbx2 = mesh.nmbx3 = nb;
  mesh.bc_x1 = DK_X1_PERIODIC; mesh.bc_x3 = DK_X3_PERIODIC;
  std::vector<dk_block> blocks(mesh.nmb);
  auto edge = [&](int l, bool hi) {
    if (!hi) return (l == 0) ? 0.0 : LeftEdge(l, nb, 0.0, 1.0);
    return (l == nb - 1) ? 1.0 : LeftEdge(l + 1, nb, 0.0, 1.0);
  };
  for (int bz = 0; bz < nb; ++bz)
    for (int by = 0; by < nb; ++by)
      for (int bx = 0; bx < nb; ++bx) {
        dk_block &b = blocks[bx + nb*(by + nb*bz)];
        b.x1min = edge(bx, false); b.x1max = edge(bx, true);
        b.x2min = edge(by, false); b.x2max = edge(by, true);
        b.x3min = edge(bz, false); b.x3max = edge(bz, true);
        b.lx1 = bx; b.lx2 = by; b.lx3 = bz; b.level = 0;
        for (int o3 = -1; o3 <= 1; ++o3)
          for (int o2 = -1; o2 <= 1; ++o2)
            for (int o1 = -1; o1 <= 1; ++o1) {
              const int gx = (bx + o1 + nb) % nb, gy = (by + o2 + nb) % nb,
                        gz = (bz + o3 + nb) % nb;
              b.nghbr[(o1 + 1) + 3*(o2 + 1) + 9*(o3 + 1)] = gx + nb*(gy + nb*gz);
            }
      }
  dk_physics phys{};
  phys.deposit = DK_TSC; phys.coupling = DK_IMEX; phys.solver = DK_DC1;
  phys.back_reaction = 1; phys.consistent_gas = 1;
  dk_setup su{};
  su.seed = 1; su.ppc = 2; su.nspecies = 1; su.taus[0] = 0.1; su.eps[0] = 1.0; su.rho0 = 1.0;
  su.gas_noise = 1.0e-3; su.dust_noise = 1.0e-3;
  const int64_t cap = (int64_t)su.ppc*mesh.nmb*nx*nx*nx + 16;
  const size_t ncell = (size_t)mesh.nmb*DK_NGAS*(nx + 2*ng)*(nx + 2*ng)*(nx + 2*ng);
  double *rdata = nullptr, *u0 = nullptr, *u1 = nullptr;
  int32_t *idata = nullptr;
  hipStream_t stream = nullptr;
  if (hipStreamCreate(&stream) != hipSuccess ||
      hipMalloc((void**)&rdata, sizeof(double)*DK_NRDATA*cap) != hipSuccess ||
      hipMalloc((void**)&idata, sizeof(int32_t)*DK_NIDATA*cap) != hipSuccess ||
      hipMalloc((void**)&u0, sizeof(double)*ncell) != hipSuccess ||
      hipMalloc((void**)&u1, sizeof(double)*ncell) != hipSuccess ||
      hipMemset(rdata, 0, sizeof(double)*DK_NRDATA*cap) != hipSuccess ||
      hipMemset(idata, 0, sizeof(int32_t)*DK_NIDATA*cap) != hipSuccess) {
    std::fprintf(stderr, "drive_step: device allocation failed\n");
    return 1;
  }
  dk_ctx *ctx = nullptr;
  int e;
  if ((e = dk_create(&mesh, blocks.data(), &phys, form, stream, &ctx))) return Fail("create", e);
  if ((e = dk_bind_gas(ctx, u0, u1))) return Fail("bind_gas", e);
  if ((e = dk_bind_particles(ctx, rdata, idata, cap, 0))) return Fail("bind_particles", e);
  if ((e = dk_setup_problem(ctx, &su))) return Fail("setup", e);
  if (graphs && (e = dk_set_graphs(ctx, 1))) return Fail("set_graphs", e);
  Driver drv;
  SetImEx2Plus(drv);
  Mesh pm;
  for (int n = 0; n < steps; ++n) {
    for (int s = 1; s <= drv.nexp_stages; ++s) {
      const dk_stage st = StageOf(&drv, &pm, s);
      if ((e = dk_stage_run(ctx, &st, nullptr))) return Fail("stage", e);
    }
    pm.time += pm.dt;
  }
  int32_t code = 0;
  int64_t part = -1;
  if ((e = dk_error(ctx, &code, &part))) return Fail("error", e);
  if (code) {
    std::fprintf(stderr, "drive_step: device error %s (particle %lld)\n", dk_strerror(code),
                 (long long)part);
    return 1;
  }
  double d[9];
  if ((e = dk_diag_sums(ctx, d))) return Fail("diag", e);
  if ((e = dk_export_particles(ctx))) return Fail("export", e);
  const int64_t np = dk_npart(ctx);
  std::vector<double> h((size_t)DK_NRDATA*cap);
  if (hipStreamSynchronize(stream) != hipSuccess ||
      hipMemcpy(h.data(), rdata, sizeof(double)*h.size(), hipMemcpyDeviceToHost) != hipSuccess) {
    std::fprintf(stderr, "drive_step: copy failed\n");
    return 1;
  }
  // a checksum of the exported velocities (sum over particles in store order, FP64)
  double vs = 0.0;
  for (int c : {DK_IPVX, DK_IPVY, DK_IPVZ})
    for (int64_t p = 0; p < np; ++p) vs += std::fabs(h[(size_t)c*cap + p]);
  std::printf("npart %lld graphs %lld\n", (long long)np, (long long)dk_graph_launches(ctx));
  std::printf("diag");
  for (double x : d) std::printf(" %.17g", x);
  std::printf("\nvsum %.17g\n", vs);
  dk_destroy(ctx);
  (void)hipFree(rdata); (void)hipFree(idata); (void)hipFree(u0); (void)hipFree(u1);
  (void)hipStreamDestroy(stream);
  return 0;
}
