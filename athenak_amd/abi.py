"""ctypes mirror of include/dustk.h and include/dustk_setup.h, plus the host-side pieces of the
reference driver that the C-ABI takes as plain numbers: the uniform MeshBlock table
(src/mesh/meshblock.cpp:60-120, src/coordinates/cell_locations.hpp:24-27) and the imex2+ / rk2
stage coefficients (src/driver/driver.cpp:106-118,195-203,278-300)."""

import ctypes as C

DK_NRDATA = 17
DK_NIDATA = 5
DK_NGAS = 4
(IPX, IPVX, IPY, IPVY, IPZ, IPVZ, IPX1, IPVX1, IPY1, IPVY1, IPZ1, IPVZ1,
 IPRX, IPRY, IPRZ, IPTS, IPM) = range(17)
PGID, PTAG, PSP, PLEV, PTAGH = range(5)
DK_NGP, DK_CIC, DK_TSC = 0, 1, 2
DK_IMEX, DK_PC2 = 0, 1
DK_LOCAL, DK_APPLYA, DK_DC1, DK_DC2, DK_PCG, DK_ADAPTIVE = 0, 1, 2, 3, 4, 5
DK_X1_PERIODIC, DK_X1_SHEAR = 0, 1
DK_X3_PERIODIC, DK_X3_REFLECT, DK_X3_OUTFLOW = 0, 1, 2
DK_FORM_FP64, DK_FORM_FP32 = 0, 1
DK_MAXSPEC = 8

(DK_OK, DK_EINVAL, DK_EHIP, DK_ENOMEM, DK_EHALO, DK_EROUTE, DK_ESTOP, DK_EUNSUPPORTED,
 DK_ECAPACITY, DK_ESTATE, DK_ENOTFINITE, DK_ESOLVER) = range(12)

ERRORS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "out of device memory",
          4: "particle outside deposit halo", 5: "particle routing", 6: "stopping time <= 0",
          7: "unsupported option", 8: "capacity", 9: "call order", 10: "non-finite value",
          11: "drag solver failed"}


class dk_block(C.Structure):
    _fields_ = [("x1min", C.c_double), ("x1max", C.c_double),
                ("x2min", C.c_double), ("x2max", C.c_double),
                ("x3min", C.c_double), ("x3max", C.c_double),
                ("lx1", C.c_int32), ("lx2", C.c_int32), ("lx3", C.c_int32),
                ("level", C.c_int32), ("nghbr", C.c_int32 * 27)]


class dk_mesh(C.Structure):
    _fields_ = [("nmb", C.c_int32), ("nx1", C.c_int32), ("nx2", C.c_int32),
                ("nx3", C.c_int32), ("ng", C.c_int32),
                ("x1min", C.c_double), ("x1max", C.c_double),
                ("x2min", C.c_double), ("x2max", C.c_double),
                ("x3min", C.c_double), ("x3max", C.c_double),
                ("nmbx1", C.c_int32), ("nmbx2", C.c_int32), ("nmbx3", C.c_int32),
                ("bc_x1", C.c_int32), ("bc_x3", C.c_int32), ("x3_dep_drop", C.c_int32)]


class dk_physics(C.Structure):
    _fields_ = [("deposit", C.c_int32), ("coupling", C.c_int32), ("solver", C.c_int32),
                ("back_reaction", C.c_int32), ("consistent_gas", C.c_int32),
                ("shearing_box", C.c_int32), ("stratified", C.c_int32),
                ("omega0", C.c_double), ("qshear", C.c_double),
                ("pc2_damping", C.c_int32), ("stiffness", C.c_int32), ("clouds", C.c_int32)]


class dk_stage(C.Structure):
    _fields_ = [("stage", C.c_int32), ("nexp_stages", C.c_int32),
                ("time", C.c_double), ("dt", C.c_double),
                ("gam0", C.c_double * 3), ("gam1", C.c_double * 3),
                ("beta", C.c_double * 3), ("a_twid22", C.c_double),
                ("a_impl", C.c_double)]


class dk_gas_src(C.Structure):
    _fields_ = [("shearing_box", C.c_int32), ("stratified", C.c_int32),
                ("omega0", C.c_double), ("qshear", C.c_double),
                ("const_accel", C.c_double), ("const_accel_dir", C.c_int32),
                ("hydro", C.c_int32), ("iso_cs", C.c_double), ("dfloor", C.c_double)]


class dk_output_info(C.Structure):
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("cycle", C.c_int64),
                ("file_number", C.c_int32), ("single_precision", C.c_int32),
                ("stopping_time_mode", C.c_int32), ("nspecies", C.c_int32),
                ("taus", C.c_double * 8),
                ("basename", C.c_char_p), ("output_block", C.c_char_p),
                ("source_commit", C.c_char_p), ("input_dump", C.c_char_p)]


class dk_setup(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("ppc", C.c_int32), ("nspecies", C.c_int32),
                ("taus", C.c_double * DK_MAXSPEC), ("eps", C.c_double * DK_MAXSPEC),
                ("continuous", C.c_int32),
                ("tau_lo", C.c_double * DK_MAXSPEC), ("tau_hi", C.c_double * DK_MAXSPEC),
                ("mass_slope", C.c_double), ("rho0", C.c_double),
                ("gas_vx", C.c_double), ("gas_vy", C.c_double), ("gas_vz", C.c_double),
                ("gas_noise", C.c_double), ("dust_noise", C.c_double),
                ("nsh", C.c_int32), ("omega0", C.c_double), ("qshear", C.c_double),
                ("gauss_z", C.c_int32), ("hp", C.c_double), ("gas_strat", C.c_int32),
                ("nclump", C.c_int32), ("clump_frac", C.c_double),
                ("clump_sigma", C.c_double), ("zmetal", C.c_double),
                ("block_lognormal", C.c_int32), ("block_sigma", C.c_double)]


def _left_edge(ith, n, xmin, xmax):
    """LeftEdgeX of src/coordinates/cell_locations.hpp:24-27 (same operation order)."""
    x = float(ith) / float(n)
    return (x * xmax - x * xmin) - (0.5 * xmax - 0.5 * xmin) + (0.5 * xmin + 0.5 * xmax)


def uniform_mesh(nx, nb, box, ng=2, bc_x1=DK_X1_PERIODIC, bc_x3=DK_X3_PERIODIC,
                 x3_dep_drop=0):
    """A uniform mesh of nb[0]*nb[1]*nb[2] blocks of nx[0]*nx[1]*nx[2] cells over
    box = ((x1min, x1max), (x2min, x2max), (x3min, x3max)).  Block gid = bx + nbx*(by + nby*bz);
    block edges as MeshBlock::SetPositions (src/mesh/meshblock.cpp:40-120).  Neighbours are
    periodic in x1 and x2 (the shear-periodic x1 face is also listed with its plain periodic
    neighbour; the library applies the shear), periodic in x3 unless bc_x3 is a physical face
    (then -1).  Returns (dk_mesh, array of dk_block)."""
    (x1a, x1b), (x2a, x2b), (x3a, x3b) = box
    nbx, nby, nbz = nb
    mesh = dk_mesh(nmb=nbx * nby * nbz, nx1=nx[0], nx2=nx[1], nx3=nx[2], ng=ng,
                   x1min=x1a, x1max=x1b, x2min=x2a, x2max=x2b, x3min=x3a, x3max=x3b,
                   nmbx1=nbx, nmbx2=nby, nmbx3=nbz, bc_x1=bc_x1, bc_x3=bc_x3,
                   x3_dep_drop=x3_dep_drop)
    blocks = (dk_block * mesh.nmb)()

    def edges(l, n, a, b):
        lo = a if l == 0 else _left_edge(l, n, a, b)
        hi = b if l == n - 1 else _left_edge(l + 1, n, a, b)
        return lo, hi

    for bz in range(nbz):
        for by in range(nby):
            for bx in range(nbx):
                g = bx + nbx * (by + nby * bz)
                b = blocks[g]
                b.x1min, b.x1max = edges(bx, nbx, x1a, x1b)
                b.x2min, b.x2max = edges(by, nby, x2a, x2b)
                b.x3min, b.x3max = edges(bz, nbz, x3a, x3b)
                b.lx1, b.lx2, b.lx3, b.level = bx, by, bz, 0
                for o3 in (-1, 0, 1):
                    for o2 in (-1, 0, 1):
                        for o1 in (-1, 0, 1):
                            d = (o1 + 1) + 3 * (o2 + 1) + 9 * (o3 + 1)
                            z = bz + o3
                            if bc_x3 != DK_X3_PERIODIC and (z < 0 or z >= nbz):
                                b.nghbr[d] = -1
                                continue
                            gx, gy, gz = (bx + o1) % nbx, (by + o2) % nby, z % nbz
                            b.nghbr[d] = gx + nbx * (gy + nby * gz)
    return mesh, blocks


def _edges(l, n, a, b):
    lo = a if l == 0 else _left_edge(l, n, a, b)
    hi = b if l == n - 1 else _left_edge(l + 1, n, a, b)
    return lo, hi


def smr_mesh(nx, nb, box, regions, ng=2):
    """A statically refined mesh, periodic in every direction: root blocks nb[0]*nb[1]*nb[2] of
    nx[0]*nx[1]*nx[2] cells over box, and for each level L = 1, 2, ... the region regions[L-1]
    = ((a1, b1), (a2, b2), (a3, b3)) in units of root blocks: every level-(L-1) leaf whose extent
    overlaps it with positive volume is split into its 8 children (<mesh_refinement> static
    regions of src/mesh/build_tree.cpp).  Leaves are listed depth first, children in x1-fastest
    order (the reference's Z-order of the block tree); each block's box is LeftEdgeX at its level
    (src/mesh/meshblock.cpp:40-120).  Neighbour levels may differ by at most one (checked).
    nghbr[d] holds the leaf that contains the centre of the same-level region beside the block
    in direction d when that leaf is of the block's level or coarser, else -2 (finer leaves
    there).  Returns (dk_mesh, array of dk_block)."""
    (x1a, x1b), (x2a, x2b), (x3a, x3b) = box
    nbx, nby, nbz = nb
    leaves = [(0, bx, by, bz) for bz in range(nbz) for by in range(nby) for bx in range(nbx)]
    for L, reg in enumerate(regions, start=1):
        out = []
        for (lv, l1, l2, l3) in leaves:
            s = 1.0 / (1 << lv)      # block length in root-block units
            lo, hi = (l1 * s, l2 * s, l3 * s), ((l1 + 1) * s, (l2 + 1) * s, (l3 + 1) * s)
            hit = lv == L - 1 and all(min(hi[d], reg[d][1]) > max(lo[d], reg[d][0])
                                      for d in range(3))
            if not hit:
                out.append((lv, l1, l2, l3))
                continue
            for o3 in (0, 1):
                for o2 in (0, 1):
                    for o1 in (0, 1):
                        out.append((L, 2 * l1 + o1, 2 * l2 + o2, 2 * l3 + o3))
        leaves = out
    maxlev = max(lv for lv, *_ in leaves)
    F = (nbx << maxlev, nby << maxlev, nbz << maxlev)
    slot = {}
    for g, (lv, l1, l2, l3) in enumerate(leaves):
        r = 1 << (maxlev - lv)
        for s3 in range(l3 * r, (l3 + 1) * r):
            for s2 in range(l2 * r, (l2 + 1) * r):
                for s1 in range(l1 * r, (l1 + 1) * r):
                    slot[(s1, s2, s3)] = g

    def leaf_at(f):   # f: a position in finest-slot units (periodic)
        return slot[tuple(int(f[d] // 1) % F[d] for d in range(3))]

    mesh = dk_mesh(nmb=len(leaves), nx1=nx[0], nx2=nx[1], nx3=nx[2], ng=ng,
                   x1min=x1a, x1max=x1b, x2min=x2a, x2max=x2b, x3min=x3a, x3max=x3b,
                   nmbx1=nbx, nmbx2=nby, nmbx3=nbz, bc_x1=DK_X1_PERIODIC,
                   bc_x3=DK_X3_PERIODIC, x3_dep_drop=0)
    blocks = (dk_block * mesh.nmb)()
    for g, (lv, l1, l2, l3) in enumerate(leaves):
        b = blocks[g]
        b.x1min, b.x1max = _edges(l1, nbx << lv, x1a, x1b)
        b.x2min, b.x2max = _edges(l2, nby << lv, x2a, x2b)
        b.x3min, b.x3max = _edges(l3, nbz << lv, x3a, x3b)
        b.lx1, b.lx2, b.lx3, b.level = l1, l2, l3, lv
        r = 1 << (maxlev - lv)
        for o3 in (-1, 0, 1):
            for o2 in (-1, 0, 1):
                for o1 in (-1, 0, 1):
                    d = (o1 + 1) + 3 * (o2 + 1) + 9 * (o3 + 1)
                    c = ((l1 + o1 + 0.5) * r, (l2 + o2 + 0.5) * r, (l3 + o3 + 0.5) * r)
                    n = leaf_at(c)
                    b.nghbr[d] = n if leaves[n][0] <= lv else -2
                    # 2:1 balance: the leaves just beyond this face, edge or corner (sampled
                    # at the quarter points along the directions with o = 0) differ by one level
                    # at most
                    if (o1, o2, o3) == (0, 0, 0):
                        continue
                    for f3 in (0.25, 0.75):
                        for f2 in (0.25, 0.75):
                            for f1 in (0.25, 0.75):
                                q = []
                                for l_, o_, f_ in ((l1, o1, f1), (l2, o2, f2), (l3, o3, f3)):
                                    if o_ < 0:
                                        q.append(l_ * r - 0.5)
                                    elif o_ > 0:
                                        q.append((l_ + 1) * r + 0.5)
                                    else:
                                        q.append((l_ + f_) * r)
                                if abs(leaves[leaf_at(q)][0] - lv) > 1:
                                    raise ValueError("refinement regions break the 2:1 balance "
                                                     "at block %d" % g)
    return mesh, blocks


IMEX2P_GAMMA = 1.707106781186547   # the literal of src/driver/driver.cpp:203


def imex2p_stages(time, dt, gamma=IMEX2P_GAMMA):
    """The imex2+ tableau (Driver::SetImEx2PlusCoefficients, src/driver/driver.cpp:278-300)
    as three dk_stage records (stage 3 is the identity assembly stage)."""
    out = []
    for s in range(3):
        st = dk_stage(stage=s + 1, nexp_stages=3, time=time, dt=dt)
        st.gam0[0], st.gam1[0], st.beta[0] = 1.0, 0.0, gamma
        st.gam0[1] = (2.0 * gamma - 1.0) / (2.0 * gamma * gamma)
        st.gam1[1] = (1.0 - (2.0 * gamma - 1.0) / (2.0 * gamma * gamma))
        st.beta[1] = 1.0 / (2.0 * gamma)
        st.gam0[2], st.gam1[2], st.beta[2] = 1.0, 0.0, 0.0
        st.a_twid22 = (1.0 - 2.0 * gamma * gamma) / 2.0 / gamma
        st.a_impl = gamma
        out.append(st)
    return out


def gamma_switch(dt, taus_max, coupling=DK_IMEX):
    """GammaSwitch (src/dust/dust_newdt.cpp:160-173, Krapp et al. 2024 Eq. 18): the imex2+
    diagonal coefficient 1/2 when dt exceeds the largest stopping time (stiff regime), else
    1 + 1/sqrt(2); pc2 has none (None)."""
    if coupling == DK_PC2:
        return None
    return 0.5 if dt > taus_max else IMEX2P_GAMMA


def rk2_stages(time, dt):
    """Heun / SSPRK(2,2) (src/driver/driver.cpp:106-118) as two dk_stage records."""
    out = []
    for s in range(2):
        st = dk_stage(stage=s + 1, nexp_stages=2, time=time, dt=dt)
        st.gam0[0], st.gam1[0], st.beta[0] = 0.0, 1.0, 1.0
        st.gam0[1], st.gam1[1], st.beta[1] = 0.5, 0.5, 0.5
        out.append(st)
    return out
