// Element-local volume right-hand side of the hydrostatic primitive
// equations on gfx950 (MI355X), sum-factorised.
//
// Replaces PrimitiveEquationsSolver::compute_rhs
// (src/physics/primitive_equations.cpp:428-442) and everything below it:
// update_derived (:12-29), compute_momentum_rhs (:154-199), compute_tracer_rhs
// (:201-227), compute_eta_rhs (:229-273), with the reference-coordinate
// derivatives of DG3DElementOperator::{gradient,divergence}_reference
// (src/dg/operators_3d.cpp:58-79) applied by sum factorisation instead of
// dense Kronecker mat-vecs (D_xi = D (x) I (x) I, src/dg/basis_hexahedron.cpp:321-357).
// Second derivatives are applied as D (D u) in FP64, two line contractions
// as the reference applies its Kronecker gradient twice (only D then sits in
// SGPRs), and as (D D) u with D^2 formed once on the host in FP32.
//
// Work decomposition.  A workgroup owns EPB elements per batch; every element
// is handled by n^2 threads, each owning one grid line of n nodes.  Line
// ownership rotates xi -> eta -> zeta -> xi through LDS twice per batch, once
// for the momentum terms and once for the tracers, so that seven staged
// values per node suffice (phases P1 .. P4b at the kernel).  Each value
// crosses LDS once per direction instead of once per derivative tap.
// Workgroups are persistent: while a batch is computed, the state lines of
// the workgroup's next batch are already being loaded into registers, so HBM
// traffic overlaps the FP64 arithmetic.  A workgroup holds 128, 256 or 512
// grid lines (at N = 4: 5, 10 or 20 elements per batch, 8 waves per CU in
// every case, as registers and LDS allow); wider workgroups read longer
// contiguous runs of each field plane.
//
// Index quirk reproduced on purpose: the free-surface tendency integrates
// Hu, Hv over the xi index (d % n) and differentiates the column integrals
// along eta and zeta, replicating the result along xi
// (primitive_equations.cpp:244-272).  omega is identically 0 in the RHS
// (src/solver/simulation_driver.cpp:222 -> primitive_equations.cpp:47-48), so
// zeta fluxes vanish.
#pragma once
#include <type_traits>

#include "drf_internal.hpp"
#include "interface.hpp"

namespace drf {
namespace vol {

// LDS slots per element node, used in two rounds (see the kernel).
enum : int { S_A = 0, S_B, S_C, S_D, S_E, S_F, S_G, S_NF };
static_assert(S_NF == kVolSlots, "the batch geometry counts the kernel's LDS slots");

// LW: grid lines per workgroup (128, 256 or 512), the workgroup size; the
// elements per batch EPB follow from it and from the LDS budget
// (vol_batch_elements, shared with the blocked layout's block size).
template <int N, int LW> struct Cfg {
    static constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
    static constexpr int EPB = vol_batch_elements(N, LW);
    static constexpr int NT = EPB * n2;
    static constexpr int BLOCK = ((NT + 63) / 64) * 64;
    // minimum waves per SIMD the register allocation must allow
    // (the low-storage stage carries two operand arrays through the
    // epilogue and gets one wave less)
    template <class TS, int EPI> static constexpr int wpe()
    {
        const int w = sizeof(TS) == 8 ? (N <= 4 ? 2 : 1) : (N <= 4 ? 3 : 2);
        return (EPI == EPI_LSRK && w > 1) ? w - 1 : w;
    }
};

// Operators sized for one order, passed by value (kernel-argument SGPRs).
template <int n, class TC> struct KOps {
    TC D[n * n];   // 1D derivative matrix, row-major
    TC D2[n * n];  // D * D (second derivative; read by the FP32 kernel only)
    double Dc[n * n]; // FP64 D for the free-surface column integrals
    double sw[n];     // 0.5 * LGL weights
};

template <class TS, class TC> __device__ __forceinline__ TC ld(const TS *p) { return (TC)(*p); }

// Addresses are split into a wave-uniform base (pointer + field and xi
// offsets, all from kernel arguments: SGPRs) and a 32-bit per-lane byte
// offset (element and line), so every access is one `saddr` load or store
// with a single VGPR offset.  Offsets within one field must stay below 4 GiB
// (checked on the host).
template <class TS, class TC> __device__ __forceinline__ TC ldo(const TS *ubase, uint32_t off)
{
    return (TC)(*(const TS *)((const char *)ubase + off));
}
// NT: a streaming (non-temporal) store.  The fused Runge-Kutta stages use it
// for Y and A: their stores then leave the read-only stage operands (U0)
// in the last-level cache (RK4 step 739 -> 684 us at 64x64x8, N = 4).  The
// plain RHS keeps ordinary stores (streaming ones measured 1 % slower when
// its input comes from HBM).
template <class TS, bool NT = false>
__device__ __forceinline__ void sto(TS *ubase, uint32_t off, TS v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, (TS *)((char *)ubase + off));
    else
        *(TS *)((char *)ubase + off) = v;
}
// Byte offsets of (element, line) pairs.  32-bit arithmetic is exact: the
// host checks that every offset of the launch fits in 32 bits.
//
// A view without blocks: the bathymetry and Coriolis lines (always in the
// line layout, no larger than one field of the state), and the fused face
// phase (which runs on unblocked states only).
template <class T> __device__ __forceinline__ uint32_t flat_off(const View<T> &v, int64_t e, int L)
{
    return ((uint32_t)e * (uint32_t)v.es + (uint32_t)L * (uint32_t)v.ls) * (uint32_t)sizeof(T);
}
// Element e of batch b in a state view.  All state views of one launch share
// the layout and its strides.  In the blocked layout a batch never straddles
// a block (a block's element count is a multiple of the batch's, checked at
// launch), so the block is b / (bb / EPB): a division of uniform values, once
// per batch, none per lane.  Lanes that repeat element E - 1 sit in the last
// batch, in that element's block.  Without blocks (bb = 0) the block is 0.
template <int EPB, class T>
__device__ __forceinline__ uint32_t state_off(const View<T> &v, int64_t b, int64_t e, int L)
{
    const uint32_t q = v.bb ? (uint32_t)v.bb / (uint32_t)EPB : 0xffffffffu;
    const uint32_t blk = (uint32_t)b / q;
    const uint32_t eb = (uint32_t)e - blk * (uint32_t)v.bb;
    return (blk * (uint32_t)v.bs + eb * (uint32_t)v.es + (uint32_t)L * (uint32_t)v.ls) *
           (uint32_t)sizeof(T);
}

template <class TC> __device__ __forceinline__ TC hinv_of(TC H)
{
    // PrimitiveState::update_derived, primitive_equations.cpp:23
    return (H > (TC)1e-10) ? (TC)1 / H : (TC)0;
}

// y = M x along one line (n taps), M row-major.
template <int n, class TC>
__device__ __forceinline__ void dline(const TC *__restrict__ M, const TC (&x)[n], TC (&y)[n])
{
#pragma unroll
    for (int i = 0; i < n; ++i) {
        TC s = M[i * n] * x[0];
#pragma unroll
        for (int m = 1; m < n; ++m)
            s += M[i * n + m] * x[m];
        y[i] = s;
    }
}

// Second derivative along one line.  FP64: D (D x), two contractions as the
// reference applies its gradient twice, with only D in SGPRs.  FP32: D^2
// formed once on the host (one contraction; two float matrices fit the SGPR
// file).
template <int n, class TC>
__device__ __forceinline__ void d2line(const TC *__restrict__ D, const TC *__restrict__ D2,
                                       const TC (&x)[n], TC (&y)[n])
{
    if constexpr (sizeof(TC) == 8) {
        TC t[n];
        dline<n>(D, x, t);
        dline<n>(D, t, y);
    } else {
        dline<n>(D2, x, y);
    }
}

// The state lines a thread needs for phase P1, with the Coriolis parameter
// along the line (a field, or the constant f0).
template <int n, class TC> struct Lines {
    TC hu[n], hv[n], et[n], ht[n], hs[n], hb[n], fc[n];
};

// All of P1's global loads are issued here, so that P1 itself waits only for
// them: a load issued inside P1 would make the wave wait for every store of
// the previous batch as well (gfx9 counts loads and stores in one in-order
// counter).
template <int n, int EPB, class TS, class TC>
__device__ __forceinline__ void load_lines(Lines<n, TC> &L, const View<const TS> &U,
                                           const VolGeom<TS> &geo, int64_t b, int64_t e, int line)
{
    const uint32_t uo = state_off<EPB>(U, b, e, line), ho = flat_off(geo.h, e, line);
#pragma unroll
    for (int i = 0; i < n; ++i) {
        L.hu[i] = ldo<TS, TC>(U.p + 0 * U.fs + i * U.is, uo);
        L.hv[i] = ldo<TS, TC>(U.p + 1 * U.fs + i * U.is, uo);
        L.et[i] = ldo<TS, TC>(U.p + 2 * U.fs + i * U.is, uo);
        L.ht[i] = ldo<TS, TC>(U.p + 3 * U.fs + i * U.is, uo);
        L.hs[i] = ldo<TS, TC>(U.p + 4 * U.fs + i * U.is, uo);
        L.hb[i] = ldo<TS, TC>(geo.h.p + i * geo.h.is, ho);
    }
    if (geo.f.p) {
        const uint32_t fo = flat_off(geo.f, e, line);
#pragma unroll
        for (int i = 0; i < n; ++i)
            L.fc[i] = ldo<TS, TC>(geo.f.p + i * geo.f.is, fo);
    } else {
#pragma unroll
        for (int i = 0; i < n; ++i)
            L.fc[i] = (TC)geo.f0;
    }
}

constexpr int kXcds = 8; // MI355X: 8 XCDs of 32 CUs

// Writes fields [F0, F0 + NFD) of the thread's xi-line: the tendency k (RHS)
// or a fused Runge-Kutta stage update, whose operand lines are loaded here,
// field by field, right before they are combined.  so: the line's byte
// offset in every state view (state_off).
template <int n, int F0, int NFD, class TS, class TC, int EPI>
__device__ __forceinline__ void write_fields(const Epi<TS> &epi, uint32_t so,
                                             const TC (&kv)[NFD][n])
{
    double d1 = epi.c1, d2 = epi.c2;
    if (EPI != EPI_RHS && epi.dtp) { // dt from device memory (adaptive stepping)
        const double dt = *epi.dtp;
        if (EPI == EPI_COMB)
            d2 = (d2 * dt) / epi.dtq;
        else
            d1 = (d1 * dt) / epi.dtq;
    }
    const TC c0 = (TC)epi.c0, c1 = (TC)d1, c2 = (TC)d2, c3 = (TC)epi.c3;
    if (EPI == EPI_RHS) {
        const uint32_t ro = so;
#pragma unroll
        for (int f = 0; f < NFD; ++f)
#pragma unroll
            for (int i = 0; i < n; ++i)
                sto<TS>(epi.R.p + (F0 + f) * epi.R.fs + i * epi.R.is, ro, (TS)kv[f][i]);
        return;
    }
    const View<const TS> V1 = (EPI == EPI_LSRK) ? epi.X : epi.U0;
    const View<const TS> V2 = (EPI == EPI_COMB)
                                  ? epi.X
                                  : View<const TS>{epi.A.p, epi.A.fs, epi.A.es, epi.A.ls,
                                                   epi.A.is, epi.A.bs, epi.A.bb};
    const bool two = (EPI == EPI_STAGE_ACC || EPI == EPI_RK4FIN || EPI == EPI_COMB ||
                      (EPI == EPI_LSRK && epi.acc_mode != 1));
    const uint32_t q1 = so, q2 = so, yo = so, ao = so;
#pragma unroll
    for (int f = 0; f < NFD; ++f) {
        TC o1[n], o2[n];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            o1[i] = ldo<TS, TC>(V1.p + (F0 + f) * V1.fs + i * V1.is, q1);
            o2[i] = two ? ldo<TS, TC>(V2.p + (F0 + f) * V2.fs + i * V2.is, q2) : (TC)0;
        }
#pragma unroll
        for (int i = 0; i < n; ++i) {
            const TC k = kv[f][i];
            TC y;
            if (EPI == EPI_STAGE) { // Y = U0 + c1 k ; A = k (acc_mode 1)
                if (epi.acc_mode == 1)
                    sto<TS, true>(epi.A.p + (F0 + f) * epi.A.fs + i * epi.A.is, ao, (TS)k);
                y = o1[i] + c1 * k;
            } else if (EPI == EPI_STAGE_ACC) { // A = A + c2 k ; Y = U0 + c1 k
                sto<TS, true>(epi.A.p + (F0 + f) * epi.A.fs + i * epi.A.is, ao, (TS)(o2[i] + c2 * k));
                y = o1[i] + c1 * k;
            } else if (EPI == EPI_RK4FIN) { // Y = U0 + c1 (A + k)
                y = o1[i] + c1 * (o2[i] + k);
            } else if (EPI == EPI_COMB) { // Y = c0 U0 + c1 (X + c2 k)
                y = c0 * o1[i] + c1 * (o2[i] + c2 * k);
            } else { // EPI_LSRK: A = c2 A + c1 k ; Y = X + c3 A
                const TC a = (epi.acc_mode == 1) ? c1 * k : c2 * o2[i] + c1 * k;
                sto<TS, true>(epi.A.p + (F0 + f) * epi.A.fs + i * epi.A.is, ao, (TS)a);
                y = o1[i] + c3 * a;
            }
            sto<TS, true>(epi.Y.p + (F0 + f) * epi.Y.fs + i * epi.Y.is, yo, (TS)y);
        }
    }
}

// The conforming face term of one element's xi-line, added to the line of
// the tendencies the thread has just written (the fused form of
// face_elem_kernel, faces.hip, with the same arithmetic and order: per node
// the x, y and z contributions summed, then M^-1; each interface with its own
// left/right flux order).  The thread rereads its own writes, so no barrier
// is needed.
template <class T>
__device__ __forceinline__ T face_contrib_v(int kind, const T *prm, int f, int side, T own,
                                            T other, T wq)
{
#pragma clang fp contract(off)
    const int fl = side ? (f ^ 1) : f;
    T nrm[3] = {T(0), T(0), T(0)};
    nrm[fl >> 1] = (fl & 1) ? T(1) : T(-1);
    if (side == 0)
        return T(0) + wq * scalar_flux_t<T>(kind, own, other, nrm, prm);
    return T(0) - wq * scalar_flux_t<T>(kind, other, own, nrm, prm);
}

template <int n, class TS>
__device__ __forceinline__ void face_phase(const VolFaces<TS> &fc, const View<const TS> &U,
                                           const View<TS> &R, int64_t e, int L, const TS *sWq,
                                           const TS *sMinv)
{
#pragma clang fp contract(off)
    struct Nb {
        uint32_t off;
        int side;
        bool on;
    };
    const int j = L % n, k = L / n, nq = fc.nq;
    // the faces this line lies on, and the neighbour's line across each (as
    // face_elem_kernel: x-faces at its ends, y / z when j / k is 0 or N)
    const int fy = (j == 0) ? 2 : (j == n - 1) ? 3 : -1;
    const int fz = (k == 0) ? 4 : (k == n - 1) ? 5 : -1;
    const int Ly = (n - 1 - j) + n * k, Lz = j + n * (n - 1 - k);
    auto entry = [&](int f, int line) {
        const int64_t v = f >= 0 ? fc.eface[6 * e + f] : -1;
        Nb r;
        r.on = v >= 0;
        r.side = (int)(v & 1);
        r.off = r.on ? flat_off(U, v >> 4, line) : 0u;
        return r;
    };
    const Nb x0 = entry(0, L), x1 = entry(1, L), y = entry(fy, Ly), z = entry(fz, Lz);
    const uint32_t ou = flat_off(U, e, L), orr = flat_off(R, e, L);
#pragma unroll 1
    for (int fld = 0; fld < 5; ++fld) {
        const TS *Uf = U.p + fld * U.fs;
        TS *Rf = R.p + fld * R.fs;
        TS u[n], du[n], uy[n], uz[n], ux0 = TS(0), ux1 = TS(0), acc[n];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            u[i] = ldo<TS, TS>(Uf + i * U.is, ou);
            du[i] = ldo<TS, TS>(Rf + i * R.is, orr);
            uy[i] = y.on ? ldo<TS, TS>(Uf + i * U.is, y.off) : TS(0);
            uz[i] = z.on ? ldo<TS, TS>(Uf + i * U.is, z.off) : TS(0);
            acc[i] = TS(0);
        }
        if (x0.on)
            ux0 = ldo<TS, TS>(Uf + (n - 1) * U.is, x0.off);
        if (x1.on)
            ux1 = ldo<TS, TS>(Uf, x1.off);
        if (L < nq) {
            const TS wq = sWq[L];
            if (x0.on)
                acc[0] += face_contrib_v<TS>(fc.kind, fc.p, 0, x0.side, u[0], ux0, wq);
            if (x1.on)
                acc[n - 1] += face_contrib_v<TS>(fc.kind, fc.p, 1, x1.side, u[n - 1], ux1, wq);
        }
        if (y.on) {
#pragma unroll
            for (int i = 0; i < n; ++i)
                if (i + n * k < nq)
                    acc[i] += face_contrib_v<TS>(fc.kind, fc.p, fy, y.side, u[i], uy[i],
                                                 sWq[i + n * k]);
        }
        if (z.on) {
#pragma unroll
            for (int i = 0; i < n; ++i)
                if (i + n * j < nq)
                    acc[i] += face_contrib_v<TS>(fc.kind, fc.p, fz, z.side, u[i], uz[i],
                                                 sWq[i + n * j]);
        }
#pragma unroll
        for (int i = 0; i < n; ++i)
            sto<TS>(Rf + i * R.is, orr, (TS)(du[i] + sMinv[i + n * L] * acc[i]));
    }
}

// The kernel.  Seven LDS slots per element node, six of them reused in two
// rounds (momentum, then tracers), so a workgroup of 5 N=4 elements needs
// 35 KB of staging plus 4 KB of column integrals, and the CU's occupancy is
// set by registers (4 workgroups at N=4 FP64) rather than by LDS:
//
//   P1  xi-owner   xi-derivatives of the momentum terms (registers), column
//                  integrals (FP64), stages A=Hu v, B=Hv v, C=eta, D=u, E=v, F=H,
//                  G=1/H (the dry-node guard of update_derived applied once)
//   P2a eta-owner  eta-derivatives of the momentum terms -> A, B
//   P3a zeta-owner vertical diffusion of u, v             -> A, B +=
//   P4a xi-owner   dHu, dHv complete (written out); xi-derivatives of the
//                  tracer terms; restages A=v HT, B=v HS, C=T, D=S (F, G kept)
//   P23b eta-owner eta-derivatives of the tracer fluxes    -> A, B
//        zeta-owner vertical diffusion of T, S (same phase) -> C, D
//   P4b xi-owner   dHT, dHS and d eta complete (written out)
//   P5  (FACES) xi-owner: the conforming face term added to its line
template <int N, class TS, class TC, int EPI, int LW, bool FACES>
__global__ void __launch_bounds__((Cfg<N, LW>::BLOCK)) __attribute__((amdgpu_waves_per_eu(Cfg<N, LW>::template wpe<TS, EPI>())))
    volume_kernel(const VolGeom<TS> geo, const KOps<N + 1, TC> ops, const Phys ph,
                  const View<const TS> U, const Epi<TS> epi, const int64_t nbatch)
{
    using C = Cfg<N, LW>;
    // Fused stages load their operand lines in the epilogue; the next-batch
    // prefetch is kept for the plain RHS, where registers allow it.
    constexpr bool PREFETCH = EPI == EPI_RHS;
    constexpr int n = C::n, n2 = C::n2, n3 = C::n3, EPB = C::EPB;
    __shared__ TC sm[EPB][S_NF][n3];
    // column integrals, double-buffered by batch parity: the last phase of
    // one batch reads them while the first phase of the next writes
    __shared__ double bar[2][EPB][2][n2];
    // the FP64 D of the column integrals, indexed per lane in P4b (from the
    // kernel arguments such an index compiles to a global load, and its wait
    // would drain the next batch's prefetch and this batch's stores)
    __shared__ double sDc[n * n];
    // FACES: face-rule weight by face DOF and M^-1 by node
    __shared__ TS sWq[FACES ? n2 : 1], sMinv[FACES ? n3 : 1];

    // The plain RHS runs every lane through every phase, so that no phase
    // sits behind an exec-mask branch: lanes past the tile repeat the last
    // line of the last element (in the same wavefront, in lockstep), and
    // elements past the end of the mesh repeat element E - 1 in their own LDS
    // slots.  Their stores write the values the owners write.  With the
    // stores on every path, P1 waits only for the prefetched lines and not
    // for the previous batch's stores to drain.  The fused stages, whose
    // epilogues read and rewrite their accumulators, mask such lanes.
    constexpr bool DUP = PREFETCH;
    const int tid = threadIdx.x;
    const bool in_tile = DUP || tid < C::NT;
    const int tq = (DUP && tid >= C::NT) ? C::NT - 1 : tid;
    const int le = in_tile ? tq / n2 : 0;
    const int L = in_tile ? tq - le * n2 : 0;
    const int64_t elast = geo.E - 1;
    auto elem = [&](int64_t bb) {
        const int64_t x = bb * EPB + le;
        return (DUP && x > elast) ? elast : x;
    };

    const TC g = (TC)ph.g, nu_h = (TC)ph.nu_h, nu_v = (TC)ph.nu_v, kT = (TC)ph.kT, kS = (TC)ph.kS;
    const TC *__restrict__ D = ops.D;
    const TC *__restrict__ D2 = ops.D2;
    // the FP64 D of the column integrals is D itself when TC is double:
    // reading it from D keeps one operator fewer in SGPRs
    const double *__restrict__ Dc =
        std::is_same<TC, double>::value ? (const double *)ops.D : ops.Dc;

    // XCD-aware batch order.  Workgroups are dealt round-robin to the 8 XCDs
    // (blockIdx % 8), each with its own L2.  Renumbering them so that
    // consecutive logical workgroups sit on one XCD puts neighbouring
    // batches -- whose line spans share a 128-byte cache line, the spans not
    // being line-aligned -- behind the same L2, while the grid as a whole
    // still sweeps memory in one contiguous window per round.
    const int64_t G = gridDim.x;
    const int64_t lw = (geo.xcd_remap && G % kXcds == 0)
                           ? (blockIdx.x % kXcds) * (G / kXcds) + blockIdx.x / kXcds
                           : (int64_t)blockIdx.x;
    int64_t b = lw;
    if (tid < n * n)
        sDc[tid] = Dc[tid]; // first read in P4b, after P1's barrier
    if constexpr (FACES) {
        for (int q = tid; q < n2; q += blockDim.x)
            sWq[q] = geo.faces.wq[q];
        for (int d = tid; d < n3; d += blockDim.x) {
            // 1 / (w_i w_j w_k), w = 2 sw exactly, rounded to the state type
            const TS wi = (TS)(2.0 * ops.sw[d % n]), wj = (TS)(2.0 * ops.sw[(d / n) % n]),
                     wk = (TS)(2.0 * ops.sw[d / n2]);
            const TS w = wi * wj * wk;
            sMinv[d] = TS(1) / w;
        }
    }
    Lines<n, TC> nxt;
    if (PREFETCH) {
        const int64_t e0 = elem(b);
        if (in_tile && b < nbatch && e0 < geo.E)
            load_lines<n, EPB, TS, TC>(nxt, U, geo, b, e0, L);
    }

    for (int par = 0; b < nbatch; b += G, par ^= 1) {
        const int64_t e = elem(b);
        const bool act = DUP || (in_tile && e < geo.E);
        // With the prefetch, P1 reads the prefetched lines in place (the
        // next batch's loads overwrite them only after P1's barrier), so the
        // loop carries no register copies that would wait for those loads.
        Lines<n, TC> own;
        if (!PREFETCH && act)
            load_lines<n, EPB, TS, TC>(own, U, geo, b, e, L);
        const Lines<n, TC> &cur = PREFETCH ? nxt : own;
        TC *s = &sm[le][0][0];

        // Register state carried by the xi-owner: the momentum partials
        // (P1 -> P4a), HT and HS (P1 -> P4a), the tracer partials (P4a -> P4b).
        TC pHu[n], pHv[n], ht[n], hs[n];

        // ---------------- P1: xi-line (j, k) = (L % n, L / n) ----------------
        if (act) {
            TC H[n], u[n], v[n], hinv[n];
#pragma unroll
            for (int i = 0; i < n; ++i) {
                H[i] = cur.et[i] + cur.hb[i];
                hinv[i] = hinv_of(H[i]);
                const TC hi = hinv[i];
                u[i] = cur.hu[i] * hi;
                v[i] = cur.hv[i] * hi;
                ht[i] = cur.ht[i];
                hs[i] = cur.hs[i];
            }
            // xi-derivatives (momentum_advection :275-305,
            // barotropic_pressure_gradient :322-337, horizontal_diffusion
            // :339-355) and Coriolis (:307-320)
            TC x[n], d[n];
#pragma unroll
            for (int i = 0; i < n; ++i)
                x[i] = cur.hu[i] * u[i];
            dline<n>(D, x, d);
#pragma unroll
            for (int i = 0; i < n; ++i)
                pHu[i] = -d[i] + cur.fc[i] * cur.hv[i];
#pragma unroll
            for (int i = 0; i < n; ++i)
                x[i] = cur.hv[i] * u[i];
            dline<n>(D, x, d);
#pragma unroll
            for (int i = 0; i < n; ++i)
                pHv[i] = -d[i] + (-cur.fc[i]) * cur.hu[i];
            dline<n>(D, cur.et, d);
#pragma unroll
            for (int i = 0; i < n; ++i)
                pHu[i] += (-g * H[i]) * d[i];
            d2line<n>(D, D2, u, d);
#pragma unroll
            for (int i = 0; i < n; ++i)
                pHu[i] += H[i] * (nu_h * d[i]);
            d2line<n>(D, D2, v, d);
#pragma unroll
            for (int i = 0; i < n; ++i)
                pHv[i] += H[i] * (nu_h * d[i]);
#pragma unroll
            for (int i = 0; i < n; ++i) {
                const int m = L * n + i;
                s[S_A * n3 + m] = cur.hu[i] * v[i];
                s[S_B * n3 + m] = cur.hv[i] * v[i];
                s[S_C * n3 + m] = cur.et[i];
                s[S_D * n3 + m] = u[i];
                s[S_E * n3 + m] = v[i];
                s[S_F * n3 + m] = H[i];
                s[S_G * n3 + m] = hinv[i];
            }
            // Column integrals of compute_eta_rhs (:244-254): over the xi index,
            // summed in node order from 0, in FP64 at every precision.
            double su = 0.0, sv = 0.0;
#pragma unroll
            for (int i = 0; i < n; ++i) {
                su += ops.sw[i] * (double)cur.hu[i];
                sv += ops.sw[i] * (double)cur.hv[i];
            }
            bar[par][le][0][L] = su;
            bar[par][le][1][L] = sv;
        }
        __syncthreads();

        if (PREFETCH) {
            // Next batch of this workgroup, in flight during the other phases.
            const int64_t bn = b + G;
            const int64_t en = elem(bn);
            // issued at raised wave priority, so that the SIMD's other wave
            // does not interleave its arithmetic with these loads (6 (N + 1)
            // per lane; 1 % on the FP64 and mixed RHS at 128x128x16, N = 4)
            __builtin_amdgcn_s_setprio(2);
            if (in_tile && bn < nbatch && en < geo.E)
                load_lines<n, EPB, TS, TC>(nxt, U, geo, bn, en, L);
            __builtin_amdgcn_s_setprio(0);
        }

        // ---------------- P2a: eta-line (i, k) = (L % n, L / n) ----------------
        if (act) {
            const int i0 = L % n, k0 = L / n;
            TC fuv[n], fvv[n], et[n], u[n], v[n], H[n];
#pragma unroll
            for (int j = 0; j < n; ++j) {
                const int m = i0 + n * j + n2 * k0;
                fuv[j] = s[S_A * n3 + m];
                fvv[j] = s[S_B * n3 + m];
                et[j] = s[S_C * n3 + m];
                u[j] = s[S_D * n3 + m];
                v[j] = s[S_E * n3 + m];
                H[j] = s[S_F * n3 + m];
            }
            TC d[n], aHu[n], aHv[n];
            dline<n>(D, fuv, d);
#pragma unroll
            for (int j = 0; j < n; ++j)
                aHu[j] = -d[j];
            d2line<n>(D, D2, u, d);
#pragma unroll
            for (int j = 0; j < n; ++j)
                aHu[j] += H[j] * (nu_h * d[j]);
            dline<n>(D, fvv, d);
#pragma unroll
            for (int j = 0; j < n; ++j)
                aHv[j] = -d[j];
            dline<n>(D, et, d);
#pragma unroll
            for (int j = 0; j < n; ++j)
                aHv[j] += (-g * H[j]) * d[j];
            d2line<n>(D, D2, v, d);
#pragma unroll
            for (int j = 0; j < n; ++j)
                aHv[j] += H[j] * (nu_h * d[j]);
#pragma unroll
            for (int j = 0; j < n; ++j) {
                const int m = i0 + n * j + n2 * k0;
                s[S_A * n3 + m] = aHu[j];
                s[S_B * n3 + m] = aHv[j];
            }
        }
        __syncthreads();

        // ---------------- P3a: zeta-line (i, j) = L, momentum ----------------
        if (act) {
            TC u[n], v[n], hi[n];
#pragma unroll
            for (int k = 0; k < n; ++k) {
                const int m = L + n2 * k;
                u[k] = s[S_D * n3 + m];
                v[k] = s[S_E * n3 + m];
                hi[k] = s[S_G * n3 + m];
            }
            // vertical_diffusion :357-375: nu_v * d2/dzeta2 * Hinv
            TC d[n];
            d2line<n>(D, D2, u, d);
#pragma unroll
            for (int k = 0; k < n; ++k)
                s[S_A * n3 + L + n2 * k] += nu_v * d[k] * hi[k];
            d2line<n>(D, D2, v, d);
#pragma unroll
            for (int k = 0; k < n; ++k)
                s[S_B * n3 + L + n2 * k] += nu_v * d[k] * hi[k];
        }
        __syncthreads();

        // ---------------- P4a: xi-line again: momentum done, tracers staged ----------------
        TC pHT[n], pHS[n];
        if (act) {
            TC kv[2][n];
#pragma unroll
            for (int i = 0; i < n; ++i) {
                const int m = L * n + i;
                kv[0][i] = pHu[i] + s[S_A * n3 + m];
                kv[1][i] = pHv[i] + s[S_B * n3 + m];
            }
            write_fields<n, 0, 2, TS, TC, EPI>(epi, state_off<EPB>(U, b, e, L), kv);
            TC u[n], v[n], hi[n];
#pragma unroll
            for (int i = 0; i < n; ++i) {
                const int m = L * n + i;
                u[i] = s[S_D * n3 + m];
                v[i] = s[S_E * n3 + m];
                hi[i] = s[S_G * n3 + m];
            }
            // tracer_advection :377-393 along xi
            TC x[n], d[n];
#pragma unroll
            for (int i = 0; i < n; ++i)
                x[i] = u[i] * ht[i];
            dline<n>(D, x, d);
#pragma unroll
            for (int i = 0; i < n; ++i)
                pHT[i] = -d[i];
#pragma unroll
            for (int i = 0; i < n; ++i)
                x[i] = u[i] * hs[i];
            dline<n>(D, x, d);
#pragma unroll
            for (int i = 0; i < n; ++i)
                pHS[i] = -d[i];
#pragma unroll
            for (int i = 0; i < n; ++i) {
                const int m = L * n + i;
                s[S_A * n3 + m] = v[i] * ht[i];
                s[S_B * n3 + m] = v[i] * hs[i];
                s[S_C * n3 + m] = ht[i] * hi[i];
                s[S_D * n3 + m] = hs[i] * hi[i];
            }
        }
        __syncthreads();

        // ---------------- P23b: eta-line and zeta-line, tracers ----------------
        // Both roles in one phase: the eta-owner reads and rewrites A, B along
        // its eta-line, the zeta-owner C, D along its zeta-line, so no two
        // threads touch the same node of the same slot.
        if (act) {
            const int i0 = L % n, k0 = L / n;
            TC ftv[n], fsv[n];
#pragma unroll
            for (int j = 0; j < n; ++j) {
                const int m = i0 + n * j + n2 * k0;
                ftv[j] = s[S_A * n3 + m];
                fsv[j] = s[S_B * n3 + m];
            }
            TC dT[n], dS[n];
            dline<n>(D, ftv, dT);
            dline<n>(D, fsv, dS);
#pragma unroll
            for (int j = 0; j < n; ++j) {
                const int m = i0 + n * j + n2 * k0;
                s[S_A * n3 + m] = -dT[j];
                s[S_B * n3 + m] = -dS[j];
            }
            TC T[n], S[n], hi[n];
#pragma unroll
            for (int k = 0; k < n; ++k) {
                const int m = L + n2 * k;
                T[k] = s[S_C * n3 + m];
                S[k] = s[S_D * n3 + m];
                hi[k] = s[S_G * n3 + m];
            }
            TC d[n];
            d2line<n>(D, D2, T, d);
#pragma unroll
            for (int k = 0; k < n; ++k)
                s[S_C * n3 + L + n2 * k] = kT * d[k] * hi[k];
            d2line<n>(D, D2, S, d);
#pragma unroll
            for (int k = 0; k < n; ++k)
                s[S_D * n3 + L + n2 * k] = kS * d[k] * hi[k];
        }
        __syncthreads();

        // ---------------- P4b: xi-line: free surface and tracers done ----------------
        if (act) {
            const int j0 = L % n, k0 = L / n;
            // eta tendency (:256-272): -(D_x2d HU_bar + D_y2d HV_bar), where the
            // 2D "x" index is j and the 2D "y" index is k.
            double s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int m = 0; m < n; ++m) {
                s1 += sDc[j0 * n + m] * bar[par][le][0][m + n * k0];
                s2 += sDc[k0 * n + m] * bar[par][le][1][j0 + n * m];
            }
            const TC deta = (TC)(-(s1 + s2));
            TC kv[3][n];
#pragma unroll
            for (int i = 0; i < n; ++i) {
                const int m = L * n + i;
                kv[0][i] = deta;
                kv[1][i] = (pHT[i] + s[S_A * n3 + m]) + s[S_C * n3 + m];
                kv[2][i] = (pHS[i] + s[S_B * n3 + m]) + s[S_D * n3 + m];
            }
            write_fields<n, 2, 3, TS, TC, EPI>(epi, state_off<EPB>(U, b, e, L), kv);
        }
        // ---------------- P5 (FACES): the conforming face term ----------------
        // Not by the lanes that repeat element E - 1 from another wavefront:
        // their rereads could meet the owner's final values.
        if constexpr (FACES) {
            if (b * EPB + le < geo.E)
                face_phase<n, TS>(geo.faces, U, epi.R, e, L, sWq, sMinv);
        }
        // No barrier here: this phase reads only its own line's slots and the
        // column integrals of this batch's parity; the next batch's P1 writes
        // its own line's slots and the other parity.
    }
}

template <int n, class TC> KOps<n, TC> make_ops(const Ops &o)
{
    KOps<n, TC> r{};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int m = 0; m < n; ++m)
                s += o.D[i * n + m] * o.D[m * n + j];
            r.D[i * n + j] = (TC)o.D[i * n + j];
            r.D2[i * n + j] = (TC)s;
            r.Dc[i * n + j] = o.D[i * n + j];
        }
    for (int i = 0; i < n; ++i)
        r.sw[i] = o.sw[i];
    return r;
}

inline int num_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            cus = p.multiProcessorCount;
        if (cus <= 0)
            cus = 256;
    }
    return cus;
}

template <int N, class TS, class TC, int EPI, int LW, bool FACES>
hipError_t launch_ne(const VolGeom<TS> &geo, const Ops &ops, const Phys &ph, View<const TS> U,
                     const Epi<TS> &epi, hipStream_t s)
{
    using C = Cfg<N, LW>;
    if (geo.E <= 0)
        return hipSuccess;
    // blocked states: whole batches per block; the face phase reads unblocked states only
    if (U.bb && (U.bb % C::EPB != 0 || FACES))
        return hipErrorInvalidValue;
    const int64_t nbatch = (geo.E + C::EPB - 1) / C::EPB;
    auto kern = volume_kernel<N, TS, TC, EPI, LW, FACES>;
    // resident workgroups per CU (queried once per instantiation: the launch
    // path stays free of runtime queries)
    static const int per_cu = [&] {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kern, C::BLOCK, 0) != hipSuccess ||
            v < 1)
            v = 1;
        return v;
    }();
    // persistent workgroups: a few rounds per workgroup so the next batch's
    // loads overlap the current batch's arithmetic
    const int wpc = (geo.wg_per_cu > 0 && geo.wg_per_cu < per_cu) ? geo.wg_per_cu : per_cu;
    if (EPI == EPI_RHS) // what a requested shape amounts to, for the autotuner's report
        g_rhs_per_cu[sizeof(TS) == 8 ? 0 : 1][LW == 128 ? 0 : LW == 256 ? 1 : 2][N][FACES] =
            per_cu;
    int64_t grid = (int64_t)num_cus() * wpc;
    if (grid > nbatch)
        grid = nbatch;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(C::BLOCK), 0, s, geo,
                       make_ops<N + 1, TC>(ops), ph, U, epi, nbatch);
    return hipGetLastError();
}

template <int N, class TS, class TC, int LW>
hipError_t launch_n(const VolGeom<TS> &geo, const Ops &ops, const Phys &ph, View<const TS> U,
                    const Epi<TS> &epi, hipStream_t s)
{
    if constexpr (N <= 4) // face plans are built for orders 1-4
        if (epi.mode == EPI_RHS && geo.faces.eface)
            return launch_ne<N, TS, TC, EPI_RHS, LW, true>(geo, ops, ph, U, epi, s);
    switch (epi.mode) {
    case EPI_RHS: return launch_ne<N, TS, TC, EPI_RHS, LW, false>(geo, ops, ph, U, epi, s);
    case EPI_STAGE: return launch_ne<N, TS, TC, EPI_STAGE, LW, false>(geo, ops, ph, U, epi, s);
    case EPI_STAGE_ACC: return launch_ne<N, TS, TC, EPI_STAGE_ACC, LW, false>(geo, ops, ph, U, epi, s);
    case EPI_RK4FIN: return launch_ne<N, TS, TC, EPI_RK4FIN, LW, false>(geo, ops, ph, U, epi, s);
    case EPI_COMB: return launch_ne<N, TS, TC, EPI_COMB, LW, false>(geo, ops, ph, U, epi, s);
    case EPI_LSRK: return launch_ne<N, TS, TC, EPI_LSRK, LW, false>(geo, ops, ph, U, epi, s);
    default: return hipErrorInvalidValue;
    }
}

// Orders above 4 run at 128 lines per workgroup only: their registers exceed
// what wider workgroups allow (256 VGPRs at two waves per SIMD).
template <class TS, class TC, int LW>
hipError_t dispatch(int order, const VolGeom<TS> &geo, const Ops &ops, const Phys &ph,
                    View<const TS> U, const Epi<TS> &epi, hipStream_t s)
{
    switch (order) {
    case 1: return launch_n<1, TS, TC, LW>(geo, ops, ph, U, epi, s);
    case 2: return launch_n<2, TS, TC, LW>(geo, ops, ph, U, epi, s);
    case 3: return launch_n<3, TS, TC, LW>(geo, ops, ph, U, epi, s);
    case 4: return launch_n<4, TS, TC, LW>(geo, ops, ph, U, epi, s);
    default: break;
    }
    if constexpr (LW == 128) {
        switch (order) {
        case 5: return launch_n<5, TS, TC, LW>(geo, ops, ph, U, epi, s);
        case 6: return launch_n<6, TS, TC, LW>(geo, ops, ph, U, epi, s);
        case 7: return launch_n<7, TS, TC, LW>(geo, ops, ph, U, epi, s);
        case 8: return launch_n<8, TS, TC, LW>(geo, ops, ph, U, epi, s);
        default: break;
        }
    }
    return hipErrorInvalidValue;
}

} // namespace vol
} // namespace drf
