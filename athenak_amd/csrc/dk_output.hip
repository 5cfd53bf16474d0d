// dk_output.hip -- the dust science outputs of one rank: the particle file (pbin, format
// version 2) and the particle history row (phst) with its file.
//
// pbin: ParticleBinaryOutput (src/outputs/pbin.cpp:6-60 the format, :450-1026 the writer).
// phst: DustGasDrag::ParticleHistory (src/dust/dust_science.cpp:372-484) and
// ParticleHistoryOutput::WriteOutputFile (src/outputs/phst.cpp:80-163).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>
#include <iomanip>
#include <string>
#include <vector>

#include "dk_internal.hpp"

using dk::DevBlock;
using dk::Geo;

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 4096;          // particles per partial sum (dust_science.cpp:398)
constexpr int kSums = 7;

#define PR(c, p) r[(long long)(c)*cap + (p)]
#define PI(c, p) iv[(long long)(c)*cap + (p)]

// the particle sums of the history row, one chunk of kChunk particles per thread in index
// order: count, mass, m v (3), m |v|^2/2, m z^2 over the rows of non-zero mass
// (dust_science.cpp:398-417)
__global__ void k_phst_chunks(const double *r, long long cap, long long np, int nchunk,
                              double *part) {
  const int c = blockIdx.x*blockDim.x + threadIdx.x;
  if (c >= nchunk) return;
  double a[kSums] = {0, 0, 0, 0, 0, 0, 0};
  const long long p0 = (long long)c*kChunk;
  const long long p1 = (p0 + kChunk < np) ? p0 + kChunk : np;
  for (long long p=p0; p<p1; ++p) {
    const double m = PR(DK_IPM,p);
    if (m == 0.0) continue;
    const double v1 = PR(DK_IPVX,p), v2 = PR(DK_IPVY,p), v3 = PR(DK_IPVZ,p), z = PR(DK_IPZ,p);
    a[0] += 1.0;
    a[1] += m;
    a[2] += m*v1; a[3] += m*v2; a[4] += m*v3;
    a[5] += 0.5*m*(v1*v1 + v2*v2 + v3*v3);
    a[6] += m*z*z;
  }
  for (int q=0; q<kSums; ++q) part[(long long)c*kSums + q] = a[q];
}

// the mass density of the point particles with the coupling's stencil on the owner block
// (DepositPointScienceFields, src/dust/dust_gravity_points.cpp:273-301, mass only); the ghost
// parts are folded afterwards.  Refinable clouds of level 0 on a uniform mesh deposit the hat's
// exact cell integrals, which are the TSC weights (the representation's "uniform level-0 mesh:
// canonical TSC", dust_science.cpp:97-98): scheme = TSC for them.  bad counts rows outside
// their block's stencil.
__global__ void k_sci_mass(Geo g, const DevBlock *B, const double *r, const int32_t *iv,
                           long long cap, long long np, int scheme, double *f, int *bad) {
  for (long long p = blockIdx.x*(long long)blockDim.x + threadIdx.x; p < np;
       p += (long long)gridDim.x*blockDim.x) {
    const double mass = PR(DK_IPM,p);
    if (mass == 0.0) continue;
    const int b = PI(DK_PGID,p);
    if (!isfinite(mass) || b < 0 || b >= g.nmb || !isfinite(PR(DK_IPX,p)) ||
        !isfinite(PR(DK_IPY,p)) || !isfinite(PR(DK_IPZ,p))) {
      atomicAdd(bad, 1);
      continue;
    }
    const DevBlock &s = B[b];
    int ip, jp, kp;
    double wx[3], wy[3], wz[3];
    dk::PMWeights(PR(DK_IPX,p), s.x1min, s.x1max, g.nx1, g.is, scheme, ip, wx);
    dk::PMWeights(PR(DK_IPY,p), s.x2min, s.x2max, g.nx2, g.js, scheme, jp, wy);
    dk::PMWeights(PR(DK_IPZ,p), s.x3min, s.x3max, g.nx3, g.ks, scheme, kp, wz);
    if (ip < 1 || ip >= g.n1 - 1 || jp < 1 || jp >= g.n2 - 1 || kp < 1 || kp >= g.n3 - 1) {
      atomicAdd(bad, 1);
      continue;
    }
    const double density = mass/(s.dx1*s.dx2*s.dx3);
    for (int c=0; c<3; ++c)
      for (int j=0; j<3; ++j)
        for (int i=0; i<3; ++i) {
          const double w = density*wz[c]*wy[j]*wx[i];
          if (w == 0.0) continue;
          unsafeAtomicAdd(&f[dk::Idx(g, b, 1, 0, kp + c - 1, jp + j - 1, ip + i - 1)], w);
        }
  }
}

// the largest value over the active cells, the lowest cell index on ties
// (dust_science.cpp:430-437); per workgroup (value, index) into part
__global__ void k_sci_maxloc(Geo g, const double *f, double *part) {
  __shared__ double sv[kThreads];
  __shared__ long long sl[kThreads];
  const long long ncell = (long long)g.nx1*g.nx2*g.nx3, n = (long long)g.nmb*ncell;
  double bv = 0.0;
  long long bl = -1;
  for (long long q = blockIdx.x*(long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x*blockDim.x) {
    const int m = static_cast<int>(q/ncell);
    const long long rr = q % ncell;
    const int k = g.ks + static_cast<int>(rr/((long long)g.nx1*g.nx2));
    const int j = g.js + static_cast<int>((rr/g.nx1) % g.nx2);
    const int i = g.is + static_cast<int>(rr % g.nx1);
    const double v = f[dk::Idx(g, m, 1, 0, k, j, i)];
    if (v > bv) {bv = v; bl = q;}   // (q rises: the first of equal values stays)
  }
  sv[threadIdx.x] = bv;
  sl[threadIdx.x] = bl;
  __syncthreads();
  for (int h = kThreads/2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      const double v = sv[threadIdx.x + h];
      const long long l = sl[threadIdx.x + h];
      const double w = sv[threadIdx.x];
      const long long lw = sl[threadIdx.x];
      if (l >= 0 && (lw < 0 || v > w || (v == w && l < lw))) {
        sv[threadIdx.x] = v;
        sl[threadIdx.x] = l;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2*blockIdx.x] = sv[0];
    part[2*blockIdx.x + 1] = static_cast<double>(sl[0]);
  }
}

#undef PR
#undef PI

int Hip(hipError_t e) {
  if (e == hipSuccess) return DK_OK;
  (void)hipGetLastError();
  return (e == hipErrorOutOfMemory) ? DK_ENOMEM : DK_EHIP;
}

// the particle columns on the host: (rdata rows, idata rows) of the first n particles
int CopyColumns(dk_ctx *c, const std::vector<int> &rcols, const std::vector<int> &icols,
                std::vector<std::vector<double>> &rv, std::vector<std::vector<int32_t>> &iv) {
  const long long n = c->np;
  rv.assign(rcols.size(), std::vector<double>(n));
  iv.assign(icols.size(), std::vector<int32_t>(n));
  int e = DK_OK;
  for (size_t q = 0; q < rcols.size() && !e && n > 0; ++q)
    e = Hip(hipMemcpyAsync(rv[q].data(), c->rdata + (long long)rcols[q]*c->cap,
                           sizeof(double)*n, hipMemcpyDeviceToHost, c->stream));
  for (size_t q = 0; q < icols.size() && !e && n > 0; ++q)
    e = Hip(hipMemcpyAsync(iv[q].data(), c->idata + (long long)icols[q]*c->cap,
                           sizeof(int32_t)*n, hipMemcpyDeviceToHost, c->stream));
  if (!e) e = Hip(hipStreamSynchronize(c->stream));
  return e;
}

const char *Str(const char *s, const char *dflt) { return (s != nullptr) ? s : dflt; }

const char *StencilName(int d) { return (d == DK_TSC) ? "tsc" : (d == DK_CIC) ? "cic" : "ngp"; }

const char *ModeName(int m) {
  return (m == 0) ? "species_fixed" : (m == 1) ? "particle_static" : "dynamic";
}

// pbin.cpp:154-172: FNV-1a 64 over (level, lx1, lx2, lx3) of every block in gid order as
// little-endian int64
uint64_t TopologyHash(const dk_ctx *c) {
  uint64_t h = 1469598103934665603ULL;
  for (const dk_block &b : c->B) {
    const int64_t v[4] = {b.level, b.lx1, b.lx2, b.lx3};
    for (int q=0; q<4; ++q)
      for (int k=0; k<8; ++k) {
        h ^= (static_cast<uint64_t>(v[q]) >> (8*k)) & 0xffu;
        h *= 1099511628211ULL;
      }
  }
  return h;
}

uint64_t Align64(uint64_t a) { return a + (64 - a % 64) % 64; }

struct Col {
  std::string table, name, dtype, meaning;
  uint64_t itemsize, count, offset, bytes;
};

}  // namespace

int dk_write_pbin(dk_ctx *c, const char *path, const dk_output_info *info) {
  if (c == nullptr || path == nullptr || info == nullptr || c->rdata == nullptr)
    return c == nullptr || path == nullptr || info == nullptr ? DK_EINVAL : DK_ESTATE;
  if (info->nspecies < 1 || info->nspecies > DK_MAXSPEC) return DK_EINVAL;
  int e = dk_export_particles(c);
  if (e) return e;
  // ---- the active rows (pbin.cpp:450-485)
  std::vector<std::vector<double>> rv;
  std::vector<std::vector<int32_t>> iv;
  const std::vector<int> rc = {DK_IPX, DK_IPY, DK_IPZ, DK_IPVX, DK_IPVY, DK_IPVZ, DK_IPM,
                               DK_IPTS};
  const std::vector<int> ic = {DK_PTAG, DK_PTAGH, DK_PSP, DK_PLEV};
  if ((e = CopyColumns(c, rc, ic, rv, iv))) return e;
  const long long np = c->np;
  std::vector<long long> act;
  long long nzero = 0, nbad = 0, nneg = 0;
  for (long long p=0; p<np; ++p) {
    const double m = rv[6][p];
    if (m == 0.0) {++nzero; continue;}
    act.push_back(p);
    if (!(m > 0.0) || !std::isfinite(m)) ++nbad;
    if (iv[1][p] < 0) ++nneg;
  }
  const uint64_t ntotal = act.size();
  const bool clouds = c->P.clouds != 0;
  // ---- constants (pbin.cpp:534-568)
  bool constant_sp_ts = (info->nspecies == 1 && info->stopping_time_mode == 0);
  long long ndiffer = 0;
  const double ts0 = info->taus[0];
  if (constant_sp_ts) {
    for (long long p : act)
      if (iv[2][p] != 0 || rv[7][p] != ts0) ++ndiffer;
    if (ndiffer > 0) constant_sp_ts = false;
  }
  double mlo = std::numeric_limits<double>::max(), mhi = -std::numeric_limits<double>::max();
  for (long long p : act) {mlo = std::fmin(mlo, rv[6][p]); mhi = std::fmax(mhi, rv[6][p]);}
  const bool constant_mass = (ntotal > 0 && nbad == 0 && mlo == mhi);
  const bool single = info->single_precision != 0;
  const char *rdt = single ? "f4" : "f8";
  const uint64_t rsz = single ? 4 : 8;
  const uint64_t nblocks = c->B.size();
  // ---- the columns (pbin.cpp:581-626), little-endian
  std::vector<Col> cols;
  auto col = [&](const char *t, const char *n, const char *dt, uint64_t size,
                 const std::string &meaning) {
    cols.push_back(Col{t, n, std::string("<") + dt, meaning, size,
                       std::string(t) == "particles" ? ntotal : nblocks, 0, 0});
  };
  col("particles", "id", "u8", 8, "numerical sample identity: the 64-bit tag; split / merge / "
      "resampling create and remove samples, an id is not a grain lineage");
  const char *species_meaning = "species_id (0-based; input index N = species_id + 1)";
  const char *ts_meaning = "stopping time t_s (code time)";
  const char *mass_meaning = "macro-particle mass (code units; not a grain mass)";
  if (!constant_sp_ts) col("particles", "species", "u4", 4, species_meaning);
  col("particles", "x1", rdt, rsz, "position x1 (code length)");
  col("particles", "x2", rdt, rsz, "position x2 (code length)");
  col("particles", "x3", rdt, rsz, "position x3 (code length)");
  col("particles", "v1", rdt, rsz, "velocity component 1 (see [frame])");
  col("particles", "v2", rdt, rsz, "velocity component 2 (see [frame])");
  col("particles", "v3", rdt, rsz, "velocity component 3 (see [frame])");
  if (!constant_mass) col("particles", "mass", rdt, rsz, mass_meaning);
  if (!constant_sp_ts) col("particles", "stopping_time", rdt, rsz, ts_meaning);
  if (clouds)
    col("particles", "shape", "u2", 2, "cloud shape code v1: bits 0-3 cloud level, bits 4-5 "
        "lobe along x1, 6-7 lobe along x2, 8-9 lobe along x3 (0 none, 1 lower, 2 upper, "
        "3 both)");
  col("blocks", "gid", "i8", 8, "MeshBlock global id");
  col("blocks", "rank", "i4", 4, "MPI rank that wrote the block");
  col("blocks", "level", "i4", 4, "refinement level above the root level");
  col("blocks", "lx1", "i8", 8, "logical location x1 (at the block's level)");
  col("blocks", "lx2", "i8", 8, "logical location x2");
  col("blocks", "lx3", "i8", 8, "logical location x3");
  col("blocks", "x1min", "f8", 8, "block box x1min");
  col("blocks", "x1max", "f8", 8, "block box x1max");
  col("blocks", "x2min", "f8", 8, "block box x2min");
  col("blocks", "x2max", "f8", 8, "block box x2max");
  col("blocks", "x3min", "f8", 8, "block box x3min");
  col("blocks", "x3max", "f8", 8, "block box x3max");
  uint64_t at = 0;
  for (Col &k : cols) {
    k.offset = at;
    k.bytes = k.count*k.itemsize;
    at = Align64(at + k.bytes);
  }
  const uint64_t data_bytes = at;
  // ---- the text header (pbin.cpp:700-900)
  const dk_mesh &M = c->M;
  int maxlev = 0;
  for (const dk_block &b : c->B) maxlev = std::max(maxlev, static_cast<int>(b.level));
  std::ostringstream h;
  h << std::setprecision(17);
  h << "AthenaK particle science file (pbin) format_version = 2\n";
  h << "[file]\nformat = AKPBIN\nformat_version = 2\nbyte_order = little\n";
  h << "n_particles = " << ntotal << "\nn_blocks = " << nblocks << "\n";
  h << "written_by_ranks = 1\n";
  h << "completeness = footer AKPBEND1 + file size; written as .tmp, renamed when complete\n";
  h << "[snapshot]\n";
  h << "time = " << info->time << "\ncycle = " << info->cycle << "\n";
  h << "dt = " << info->dt << "\n";
  h << "basename = " << Str(info->basename, "") << "\n";
  h << "output_block = " << Str(info->output_block, "") << "\n";
  h << "file_number = " << info->file_number << "\n";
  h << "source_commit = " << Str(info->source_commit, "unknown") << "\n";
  h << "mesh_update_seq = 0\n";
  {
    std::ostringstream hx;
    hx << "0x" << std::hex << std::setw(16) << std::setfill('0') << TopologyHash(c);
    h << "block_topology_hash = " << hx.str() << "  # FNV-1a 64 of (level, lx1, lx2, lx3) as "
      << "int64 for every block in gid order\n";
  }
  h << "output_point = end of a cycle, before that cycle's mesh refinement check (the first "
    << "output: after the problem generator; the last: after the final check)\n";
  h << "[mesh]\nndim = 3\n";
  h << "nx1 = " << M.nx1*M.nmbx1 << "\nnx2 = " << M.nx2*M.nmbx2 << "\nnx3 = " << M.nx3*M.nmbx3
    << "\n";
  h << "x1min = " << M.x1min << "\nx1max = " << M.x1max << "\n";
  h << "x2min = " << M.x2min << "\nx2max = " << M.x2max << "\n";
  h << "x3min = " << M.x3min << "\nx3max = " << M.x3max << "\n";
  h << "meshblock_nx1 = " << M.nx1 << "\nmeshblock_nx2 = " << M.nx2 << "\nmeshblock_nx3 = "
    << M.nx3 << "\n";
  h << "levels = " << maxlev << "  # finest level above the root\n";
  h << "multilevel = " << (c->smr ? 1 : 0) << "\nadaptive = 0\n";
  {
    const char *x1 = (M.bc_x1 == DK_X1_SHEAR) ? "shear_periodic" : "periodic";
    const char *x3 = (M.bc_x3 == DK_X3_REFLECT) ? "reflect" :
                     (M.bc_x3 == DK_X3_OUTFLOW) ? "outflow" : "periodic";
    h << "bc_ix1 = " << x1 << "\nbc_ox1 = " << x1 << "\nbc_ix2 = periodic\nbc_ox2 = periodic\n"
      << "bc_ix3 = " << x3 << "\nbc_ox3 = " << x3 << "\n";
  }
  h << "[particles]\n";
  h << "rows = active rows only (the capacity tail and zero-mass rows are left out)\n";
  h << "row_order = by writing rank, then the rank's array order (unsorted; neither the order "
    << "nor the rank layout is stable between outputs)\n";
  h << "zero_mass_rows_excluded = " << nzero << "\n";
  h << "invalid_mass_rows = " << nbad << "  # negative or non-finite mass (written)\n";
  h << "negative_tag_rows = " << nneg << "  # tags >= 2^63 (high word below zero; never "
    << "issued; the key keeps its version-1 name)\n";
  if (clouds) h << "shape_unknown_bit_rows = 0\n";
  h << "owner = not stored: the MeshBlock whose half-open box [min, max) holds the position "
    << "(blocks table)\n";
  {
    std::string names;
    if (constant_sp_ts) names += "species stopping_time";
    if (constant_mass) names += std::string(names.empty() ? "" : " ") + "mass";
    h << "constant_columns = " << (names.empty() ? std::string("none") : names);
    if (!names.empty()) h << "  # the same value in every row, given in [constants]";
    if (!constant_sp_ts && ndiffer > 0)
      h << (names.empty() ? "  # " : "; ") << "one species, species_fixed, but " << ndiffer
        << " rows differ from species 0 / its tau_s: kept as columns";
    h << "\n";
  }
  h << "particle_real_precision = " << (single ? "single" : "double")
    << "  # positions, velocities, mass, stopping time (<outputN>/precision)\n";
  h << "[species]\n";
  h << "nspecies = " << info->nspecies << "\n";
  h << "stopping_time_mode = " << ModeName(info->stopping_time_mode) << "  # stopping_time "
    << "(column or constant) is the value each particle uses\n";
  h << "back_reaction = " << (c->P.back_reaction ? 1 : 0) << "\n";
  for (int s=0; s<info->nspecies; ++s)
    h << "species." << s << " = input_index " << (s + 1) << ", tau_s " << info->taus[s] << "\n";
  // the representation (DustGasDrag::ScienceRepresentation, dust_science.cpp:72-116)
  h << "[representation]\n";
  h << "representation = " << (clouds ? "refinable_hat" : std::string(StencilName(
                                        c->P.deposit)) + "_owner") << "\n";
  h << "representation_version = 1\n";
  h << "root_dx1 = " << (M.x1max - M.x1min)/(M.nx1*M.nmbx1) << "\nroot_dx2 = "
    << (M.x2max - M.x2min)/(M.nx2*M.nmbx2) << "\nroot_dx3 = "
    << (M.x3max - M.x3min)/(M.nx3*M.nmbx3) << "\n";
  h << "mesh_levels = " << maxlev << "\n";
  if (clouds) {
    h << "cloud_kernel = hat  # S(s) = max(0, 1 - |s|), s = (x - x_p)/ell, per direction\n";
    h << "cloud_half_width = root_dx_d * 2^-level  # ell_d of a cloud of level `level`; the "
      << "same physical width on any leaf\n";
    h << "cloud_lobe_code = 0 none, 1 lower, 2 upper, 3 both  # per direction\n";
    h << "cloud_lobe_shape = lobe 1/2: (2/3) S_ell(x - x_p) + (1/3) S_ell(x - x_p -/+ ell); "
      << "lobe 3: hat of half-width 2 ell\n";
    h << "cloud_support = lobe 0 [x_p - ell, x_p + ell], 1 [x_p - 2 ell, x_p + ell], 2 [x_p - "
      << "ell, x_p + 2 ell], 3 [x_p - 2 ell, x_p + 2 ell]\n";
    h << "cloud_max_level = " << maxlev << "\n";
    h << "cloud_dims = 3  # directions with a cloud shape (a 2-D run has none along x3)\n";
    h << "mesh_deposit = exact integral of the cloud shape over the active leaf cells (uniform "
      << "level-0 mesh: canonical TSC)\n";
  } else {
    h << "stencil = " << StencilName(c->P.deposit) << "  # on the cells of the owner MeshBlock "
      << "(the block whose half-open box holds x_p)\n";
    h << "transfer = none  # instant_switch: deposits in ghost cells are added to the active "
      << "leaf cells they overlap, by volume\n";
    h << "mesh_deposit = owner-block stencil, ghost-cell deposits remapped conservatively\n";
  }
  if (M.bc_x1 == DK_X1_SHEAR)
    h << "x1_faces = shear_periodic  # a point beyond the outer x1 face is (x - Lx, y + q Omega "
      << "Lx t mod Ly)\n";
  if (M.bc_x3 != DK_X3_PERIODIC)
    h << "x3_walls = fold  # the part of a deposit beyond a physical x3 face is mirrored back "
      << "(mass conserving)\n";
  h << "[frame]\npositions = lab frame, code length\n";
  if (c->P.shearing_box) {
    h << "velocity_frame = shear_relative  # velocities relative to the background shear -q "
      << "Omega x1\n";
    h << "shear_q = " << c->P.qshear << "\nshear_omega = " << c->P.omega0 << "\n";
    h << "axes = x1 radial, x2 azimuthal, x3 vertical\n";
    h << "azimuthal_velocity = v2  # v2_phys = v2 - q*Omega*x1\n";
  } else {
    h << "velocity_frame = lab\n";
  }
  h << "[columns]\n";
  h << "# column = table name dtype count offset bytes (offset from data_start)\n";
  for (const Col &k : cols)
    h << "column = " << k.table << " " << k.name << " " << k.dtype << " " << k.count << " "
      << k.offset << " " << k.bytes << "\n";
  h << "[constants]\n";
  h << "# constant = table name dtype value (the same value in every row of the table)\n";
  if (constant_sp_ts) {
    h << "constant = particles species <u4 0\n";
    h << "constant = particles stopping_time <f8 " << ts0 << "\n";
  }
  if (constant_mass) h << "constant = particles mass <f8 " << mlo << "\n";
  h << "[meaning]\n";
  for (const Col &k : cols) h << k.table << "." << k.name << " = " << k.meaning << "\n";
  if (constant_sp_ts) {
    h << "particles.species = " << species_meaning << "\n";
    h << "particles.stopping_time = " << ts_meaning << "\n";
  }
  if (constant_mass) h << "particles.mass = " << mass_meaning << "\n";
  h << "[input]\n";
  h << Str(info->input_dump, "");
  const std::string text = h.str();
  const uint64_t data_start = Align64(64 + text.size());
  const uint64_t file_bytes = data_start + data_bytes + 16;
  // ---- the file (pbin.cpp:909-986), as path.tmp renamed when complete
  std::vector<unsigned char> buf(file_bytes, 0);
  std::memcpy(buf.data(), "AKPBIN01", 8);
  const uint32_t mark = 0x01020304u, version = 2;
  std::memcpy(buf.data() + 8, &mark, 4);
  std::memcpy(buf.data() + 12, &version, 4);
  const uint64_t pre[6] = {text.size(), data_start, data_bytes, ntotal, nblocks, 0};
  std::memcpy(buf.data() + 16, pre, sizeof(pre));
  std::memcpy(buf.data() + 64, text.data(), text.size());
  for (const Col &k : cols) {
    unsigned char *dst = buf.data() + data_start + k.offset;
    if (k.table == "particles") {
      for (uint64_t q = 0; q < ntotal; ++q) {
        const long long p = act[q];
        unsigned char *d = dst + q*k.itemsize;
        if (k.name == "id") {
          const uint64_t tag = (static_cast<uint64_t>(static_cast<uint32_t>(iv[1][p])) << 32) |
                               static_cast<uint64_t>(static_cast<uint32_t>(iv[0][p]));
          std::memcpy(d, &tag, 8);
        } else if (k.name == "species") {
          const uint32_t s = static_cast<uint32_t>(iv[2][p]);
          std::memcpy(d, &s, 4);
        } else if (k.name == "shape") {
          const uint16_t s = static_cast<uint16_t>(iv[3][p] & 0xf);   // plain clouds: no lobes
          std::memcpy(d, &s, 2);
        } else {
          const int rcol = (k.name == "x1") ? 0 : (k.name == "x2") ? 1 : (k.name == "x3") ? 2 :
                           (k.name == "v1") ? 3 : (k.name == "v2") ? 4 : (k.name == "v3") ? 5 :
                           (k.name == "mass") ? 6 : 7;
          const double v = rv[rcol][p];
          if (single) {
            const float f = static_cast<float>(v);
            std::memcpy(d, &f, 4);
          } else {
            std::memcpy(d, &v, 8);
          }
        }
      }
    } else {
      for (uint64_t m = 0; m < nblocks; ++m) {
        const dk_block &b = c->B[m];
        unsigned char *d = dst + m*k.itemsize;
        if (k.dtype == "<i8") {
          const int64_t v = (k.name == "gid") ? static_cast<int64_t>(m) :
                            (k.name == "lx1") ? b.lx1 : (k.name == "lx2") ? b.lx2 : b.lx3;
          std::memcpy(d, &v, 8);
        } else if (k.dtype == "<i4") {
          const int32_t v = (k.name == "rank") ? 0 : b.level;
          std::memcpy(d, &v, 4);
        } else {
          const double v = (k.name == "x1min") ? b.x1min : (k.name == "x1max") ? b.x1max :
                           (k.name == "x2min") ? b.x2min : (k.name == "x2max") ? b.x2max :
                           (k.name == "x3min") ? b.x3min : b.x3max;
          std::memcpy(d, &v, 8);
        }
      }
    }
  }
  std::memcpy(buf.data() + data_start + data_bytes, "AKPBEND1", 8);
  std::memcpy(buf.data() + data_start + data_bytes + 8, &file_bytes, 8);
  const std::string tmp = std::string(path) + ".tmp";
  FILE *fp = std::fopen(tmp.c_str(), "wb");
  if (fp == nullptr) return DK_EINVAL;
  const bool ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
  if (std::fclose(fp) != 0 || !ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return DK_EINVAL;
  }
  return DK_OK;
}

int dk_particle_history(dk_ctx *c, double time, double stiff_thresh, double row[14]) {
  if (c == nullptr || row == nullptr) return DK_EINVAL;
  if (c->rdata == nullptr) return DK_ESTATE;
  int e = dk_export_particles(c);
  if (e) return e;
  const long long np = c->np;
  const Geo &g = c->g;
  // ---- particle sums: per chunk on the device, the chunks in order on the host
  const int nchunk = static_cast<int>((np + kChunk - 1)/kChunk);
  double sums[kSums] = {0, 0, 0, 0, 0, 0, 0};
  if (nchunk > 0) {
    double *part = nullptr;
    if ((e = Hip(hipMalloc(&part, sizeof(double)*kSums*nchunk)))) return e;
    k_phst_chunks<<<(nchunk + kThreads - 1)/kThreads, kThreads, 0, c->stream>>>(
        c->rdata, (long long)c->cap, np, nchunk, part);
    std::vector<double> ph((size_t)kSums*nchunk);
    e = Hip(hipGetLastError());
    if (!e) e = Hip(hipMemcpyAsync(ph.data(), part, sizeof(double)*ph.size(),
                                   hipMemcpyDeviceToHost, c->stream));
    if (!e) e = Hip(hipStreamSynchronize(c->stream));
    (void)hipFree(part);
    if (e) return e;
    for (int ch=0; ch<nchunk; ++ch)
      for (int q=0; q<kSums; ++q) sums[q] += ph[(size_t)ch*kSums + q];
  }
  for (int q=0; q<6; ++q) row[q] = sums[q];
  row[6] = (sums[1] > 0.0) ? std::sqrt(sums[6]/sums[1]) : 0.0;
  // ---- the densest cell of the mass deposit (point particles, and clouds on a uniform
  // mesh; clouds on a refined mesh: -1, not built)
  double where[4] = {0.0, 0.0, 0.0, -1.0};
  row[7] = -1.0;
  if (!c->P.clouds || !c->smr) {
    const long long ntot = (long long)g.nmb*g.ncb;
    double *f = nullptr, *part = nullptr;
    int *bad = nullptr;
    const int nblk = 1024;
    if ((e = Hip(hipMalloc(&f, sizeof(double)*ntot)))) return e;
    e = Hip(hipMalloc(&part, sizeof(double)*2*nblk));
    if (!e) e = Hip(hipMalloc(&bad, sizeof(int)));
    if (!e) e = Hip(hipMemsetAsync(f, 0, sizeof(double)*ntot, c->stream));
    if (!e) e = Hip(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    if (!e && np > 0) {
      const int grid = static_cast<int>(std::min<long long>((np + kThreads - 1)/kThreads, 65536));
      const int scheme = c->P.clouds ? DK_TSC : c->P.deposit;
      k_sci_mass<<<grid, kThreads, 0, c->stream>>>(g, c->d_blocks, c->rdata, c->idata,
                                                   (long long)c->cap, np, scheme, f, bad);
      e = Hip(hipGetLastError());
    }
    if (!e) e = dk::FoldT<double>(c, f, 1, time);
    int nbad = 0;
    std::vector<double> ph(2*nblk);
    if (!e) {
      k_sci_maxloc<<<nblk, kThreads, 0, c->stream>>>(g, f, part);
      e = Hip(hipGetLastError());
    }
    if (!e) e = Hip(hipMemcpyAsync(ph.data(), part, sizeof(double)*ph.size(),
                                   hipMemcpyDeviceToHost, c->stream));
    if (!e) e = Hip(hipMemcpyAsync(&nbad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (!e) e = Hip(hipStreamSynchronize(c->stream));
    (void)hipFree(f);
    (void)hipFree(part);
    (void)hipFree(bad);
    if (e) return e;
    if (nbad > 0) return DK_EHALO;
    double bv = 0.0;
    long long bl = -1;
    for (int b=0; b<nblk; ++b) {
      const double v = ph[2*b];
      const long long l = static_cast<long long>(ph[2*b + 1]);
      if (l >= 0 && (bl < 0 || v > bv || (v == bv && l < bl))) {bv = v; bl = l;}
    }
    row[7] = std::max(0.0, bv);
    if (bl >= 0 && bv > 0.0) {
      const long long ncell = (long long)g.nx1*g.nx2*g.nx3;
      const int m = static_cast<int>(bl/ncell);
      const long long rr = bl % ncell;
      const int k = static_cast<int>(rr/((long long)g.nx1*g.nx2));
      const int j = static_cast<int>((rr/g.nx1) % g.nx2);
      const int i = static_cast<int>(rr % g.nx1);
      where[0] = c->B[m].x1min + (i + 0.5)*c->dx1[m];
      where[1] = c->B[m].x2min + (j + 0.5)*c->dx2[m];
      where[2] = c->B[m].x3min + (k + 0.5)*c->dx3[m];
      where[3] = c->B[m].level;
    }
  }
  for (int q=0; q<4; ++q) row[8 + q] = where[q];
  // ---- the interval records of UpdateStiffness (chi, the rate), then reset
  double rec[2] = {-1.0, -1.0};
  if ((e = Hip(hipMemcpyAsync(rec, c->d_scalar + 14, sizeof(rec), hipMemcpyDeviceToHost,
                              c->stream))))
    return e;
  if ((e = Hip(hipStreamSynchronize(c->stream)))) return e;
  row[12] = c->P.back_reaction ? rec[0] : -1.0;
  row[13] = (c->P.back_reaction && stiff_thresh > 0.0 && rec[1] > 0.0) ? stiff_thresh/rec[1]
                                                                        : -1.0;
  const double reset[2] = {-1.0, -1.0};
  if ((e = Hip(hipMemcpyAsync(c->d_scalar + 14, reset, sizeof(reset), hipMemcpyHostToDevice,
                              c->stream))))
    return e;
  return Hip(hipStreamSynchronize(c->stream));
}

int dk_write_phst(dk_ctx *c, const char *path, const dk_output_info *info, const double row[14],
                  int32_t header) {
  if (c == nullptr || path == nullptr || info == nullptr || row == nullptr) return DK_EINVAL;
  FILE *fp = std::fopen(path, "a");
  if (fp == nullptr) return DK_EINVAL;
  static const char *labels[16] = {
      "time", "dt", "n_particles", "mass", "mom1", "mom2", "mom3", "ekin", "z_rms", "rho_max",
      "rho_max_x1", "rho_max_x2", "rho_max_x3", "rho_max_level", "chi_max", "dt_stiff_min"};
  if (header) {
    const std::string rep = c->P.clouds ? std::string("refinable_hat")
                                        : std::string(StencilName(c->P.deposit)) + "_owner";
    std::fprintf(fp, "# AthenaK dust particle history (phst) version 1\n");
    std::fprintf(fp, "# source_commit = %s; representation = %s; ndim = 3; nspecies = %d\n",
                 Str(info->source_commit, "unknown"), rep.c_str(), info->nspecies);
    std::fprintf(fp, "# species: stopping_time_mode = %s\n", ModeName(info->stopping_time_mode));
    for (int s=0; s<info->nspecies && s < DK_MAXSPEC; ++s)
      std::fprintf(fp, "# species: species.%d = input_index %d, tau_s %.17g\n", s, s + 1,
                   info->taus[s]);
    if (c->P.shearing_box) {
      std::fprintf(fp, "# velocity_frame = shear_relative: velocities relative to the "
                   "background shear -q Omega x1 (q = %.17g, Omega = %.17g); azimuthal "
                   "component v2, v2_phys = v2 - q*Omega*x1\n", c->P.qshear, c->P.omega0);
    } else {
      std::fprintf(fp, "# velocity_frame = lab\n");
    }
    std::fprintf(fp, "# a row: the state at the end of a cycle, before its mesh check (the "
                 "first row after the problem generator); sums over all ranks, active "
                 "particles only (mass != 0)\n");
    std::fprintf(fp, "# columns: 16 (the set of runs without G/R)\n");
    std::fprintf(fp, "# [1] time, [2] dt (the next step, as in .hst); [3] n_particles; [4] mass, "
                 "[5-7] mom1..3 = sum m v, [8] ekin = sum m |v|^2/2 (stored velocities), [9] "
                 "z_rms = sqrt(sum m z^2/sum m) about 0, z = x3\n");
    std::fprintf(fp, "# [10] rho_max: the largest dust density (mass per volume) on the active "
                 "cells, deposited at this time with the run's representation (the mesh field "
                 "dust_total_rho); [11-13] the center of that cell, [14] its level above the "
                 "root (rho_max = 0: center 0, level -1)\n");
    std::fprintf(fp, "# [15] chi_max: largest PC2 feedback stiffness chi = dt max(sum w m/(t_s "
                 "V rho_g)) of the cycles since the previous row (dust_stiffness_log.txt); "
                 "[16] dt_stiff_min: smallest stiff_thresh/max_rate of those cycles; both -1 "
                 "when not measured (no back-reaction, no PC2 cycle in the interval, or [16] "
                 "without stiff_thresh)\n");
    std::fprintf(fp, "# Athena++ history data\n");
    std::fprintf(fp, "#");
    for (int n=0; n<16; ++n) std::fprintf(fp, "  [%d]=%s", n + 1, labels[n]);
    std::fprintf(fp, "\n");
  }
  std::fprintf(fp, " %.16e", info->time);
  std::fprintf(fp, " %.16e", info->dt);
  for (int n=0; n<14; ++n) std::fprintf(fp, " %.16e", row[n]);
  std::fprintf(fp, "\n");
  return (std::fclose(fp) == 0) ? DK_OK : DK_EINVAL;
}
