// er_kernels.h -- gfx950 kernels of the erdst per-frame hot path.
//
// Data layout in HBM (one context, one frame in flight):
//   posn   [numatm] float4 (x,y,z as read, rotated into the upper-triangular
//          frame when the cell needs it; w = charge)
//   posu   [numatm] uint4  (lattice coordinates: the fractional coordinates
//          along the cell_normal axes times 2^32, wrapped modulo 2^32;
//          w = the charge's bits) for the pair energies
//   splw   [numatm][3][ORD] float    B-spline weights (recpcal.F90:504-538)
//   splg   [numatm] int4             base grid index per axis
//   T      [chunk][H][K2][K3] float2 spectrum after the axis-2/3 inverse DFT
//   phi    [chunk/NI][K3][K2][K1][NI] float  solute potential grids, NI
//          insertions interleaved so one wave reads 64 insertions of one
//          grid point as four 64-byte runs
//   U      [nins/NI][nummol][NI] float  uvengy, same interleave
//   C      [nins][ermax] uint32      instantaneous histograms (insdst)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "er_fft.h"

#define ER_PI_F 3.14159265358979323846f
#define ER_MAXORD 12
#define ER_BRICK 8  // PME brick side along x and y, in grid points (brick gathers)

struct FrameConst {
    double fracM[9];  // x (as read) -> fractional coordinates along the
                      // cell_normal axes, s_k = sum_j fracM[k + 3j] x_j
    double invcld[9]; // inverse of the raw cell in FP64 (the splines' fractions)
    float dsc[9];     // cell_normal / 2^32: lattice-coordinate differences -> Cartesian
    float cellN[9];  // cell_normal, column-major (realcal.F90:24)
    float invN[3];   // invcell_normal
    float lenN[3];   // cell_len_normal
    float invcl[9];  // inverse of the raw cell (engproc.F90:1049-1070)
    int is_cuboid;
    float volume;
    int rotate;      // cell not upper triangular: positions x -> Q^T x first
    float rotT[9];   // Q^T, row-major (realcal.F90:776-817)
    float invP[9];   // inverse of the cell the pair positions live in (cell_normal
                     // when rotated, else the raw cell: equal to invcl)
};

struct PairParams {
    float elecut, lwljcut, upljcut, screen;
    int ljswitch;
    int bare;  // cltype 0: plain q q / r within elecut (realcal.F90:502-506)
    // force-switch constants (realcal.F90:82-93, 899-915)
    float lw3, up3, lw6, up6;
    float repA, repB, repC, attA, attB, attC;
};

// One solute of a run, device-resident.  A run lays the insertions of all its
// solutes end to end (the run's insertion space): solute k owns insertions
// [ins_off, ins_off + nins), ins_off a multiple of the insertion interleave
// so that no tile of interleaved insertions spans two solutes; the
// insertions between nins and the next solute's ins_off are padding (zero
// spectra, never binned).  Their sites lie end to end likewise (site_off).
// The k_pairs workgroups of solute k start at pair_blk (pair_split per
// insertion) and its k_ecorr insertion slices at slice_off.  The pointers
// name the solute's own inputs and block accumulators.
struct SlotDesc {
    int nslt, ins_off, nins, site_off;
    int pair_split, pair_blk;
    int kper, nks, slice_off;
    float qslt;
    const float *sltq;             // [nslt] charges (refs)
    const int *sltlj;              // [nslt] LJ types, 0-based (refs)
    const float *wslt;             // [nins] stat_weight_solute, or null
    unsigned int *seq;             // runs of this solute (halt keys order runs first)
    double *state;                 // engnorm, engsmpl, avslf, voffset
    int *flags;
    float *fstate;
    unsigned int *minmax;          // minuv/maxuv as order keys
    double *edens, *ecorr, *eself, *slnuv;
    unsigned long long *errkey, *warnkey, *halt;
};

// the solute whose range holds x: the largest k with key(k) <= x (keys
// ascending in k; slot 0 when x lies before every range)
template <class F> __device__ __forceinline__ int slot_find(int nslot, int x, F key)
{
    int lo = 0, hi = nslot - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(mid) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
#define ER_SLOT_OF_INS(sd, nslot, n) slot_find((nslot), (n), [&](int m_) { return (sd)[m_].ins_off; })

// A pointer read from a descriptor is generic to the compiler, so its
// accesses would be flat instructions (which also count against the LDS
// wait counter).  Every descriptor pointer addresses device memory: these
// helpers access it as such.
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T *gptr(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}
template <class T> __device__ __forceinline__ void gatomic_min(T *p, T v)
{
    __hip_atomic_fetch_min(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void gatomic_max(T *p, T v)
{
    __hip_atomic_fetch_max(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int pmod(int a, int K)
{
    int r = a % K;
    return r < 0 ? r + K : r;
}

// minimum image of a difference vector (realcal.F90:271-284): the nearest
// image along each axis of a cuboid cell, else the reference's Z, Y, X
// sequence of nearest-integer shifts.  The positions are not wrapped into the
// cell first (realcal.F90:885-895): the minimum image of a difference is the
// same for every lattice image of its ends, and unwrapped FP32 positions keep
// their own rounding only.  Pair energies take their distances from lattice
// coordinates instead (lattice_d2).
__device__ __forceinline__ void min_image(const FrameConst &fc, float &dx, float &dy, float &dz)
{
    if (fc.is_cuboid) {
        dx = fmaf(-fc.lenN[0], rintf(dx * fc.invN[0]), dx);
        dy = fmaf(-fc.lenN[1], rintf(dy * fc.invN[1]), dy);
        dz = fmaf(-fc.lenN[2], rintf(dz * fc.invN[2]), dz);
    } else {
        float f = roundf(dz * fc.invN[2]);
        dx -= fc.cellN[6] * f; dy -= fc.cellN[7] * f; dz -= fc.cellN[8] * f;
        f = roundf(dy * fc.invN[1]);
        dx -= fc.cellN[3] * f; dy -= fc.cellN[4] * f; dz -= fc.cellN[5] * f;
        f = roundf(dx * fc.invN[0]);
        dx -= fc.cellN[0] * f; dy -= fc.cellN[1] * f; dz -= fc.cellN[2] * f;
    }
}

// min_image in FP64 (the solute's self energy)
__device__ __forceinline__ void min_image_d(const FrameConst &fc, double &dx, double &dy, double &dz)
{
    if (fc.is_cuboid) {
        dx = fma(-(double)fc.lenN[0], rint(dx / (double)fc.lenN[0]), dx);
        dy = fma(-(double)fc.lenN[1], rint(dy / (double)fc.lenN[1]), dy);
        dz = fma(-(double)fc.lenN[2], rint(dz / (double)fc.lenN[2]), dz);
    } else {
        double f = round(dz / (double)fc.lenN[2]);
        dx -= fc.cellN[6] * f; dy -= fc.cellN[7] * f; dz -= fc.cellN[8] * f;
        f = round(dy / (double)fc.lenN[1]);
        dx -= fc.cellN[3] * f; dy -= fc.cellN[4] * f; dz -= fc.cellN[5] * f;
        f = round(dx / (double)fc.lenN[0]);
        dx -= fc.cellN[0] * f; dy -= fc.cellN[1] * f; dz -= fc.cellN[2] * f;
    }
}

// rotate into the upper-triangular frame when the cell needs it
// (realcal.F90:814-816)
__device__ __forceinline__ void rotate_site(const FrameConst &fc, float &x, float &y, float &z)
{
    if (fc.rotate) {
        const float rx = fc.rotT[0] * x + fc.rotT[1] * y + fc.rotT[2] * z;
        const float ry = fc.rotT[3] * x + fc.rotT[4] * y + fc.rotT[5] * z;
        const float rz = fc.rotT[6] * x + fc.rotT[7] * y + fc.rotT[8] * z;
        x = rx; y = ry; z = rz;
    }
}

// Lattice coordinates of a site: its fractional coordinates along the
// cell_normal axes (from the position as read, in FP64), wrapped into
// [0, 1) and held as 32-bit fixed point.  The difference of two of them in
// 32-bit integer arithmetic is the fractional difference vector wrapped into
// [-1/2, 1/2) with no rounding at all, so the periodic wrap of a pair costs
// nothing and its distance carries only the FP32 rounding of the short
// difference itself (1e-7 relative), not that of coordinates as large as the
// box (the reference's single build, wrapping and folding FP32 coordinates,
// loses up to half an ulp of the box length per axis: 1e-5 relative at
// 2 A in an 80 A box).
__device__ __forceinline__ uint4 lattice_site(const FrameConst &fc, float x, float y, float z, float q)
{
    const double X = x, Y = y, Z = z;
    unsigned int u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double f = fc.fracM[k] * X + fc.fracM[k + 3] * Y + fc.fracM[k + 6] * Z;
        f -= floor(f);
        u[k] = (unsigned int)(unsigned long long)(long long)rint(f * 4294967296.0);  // 2^32 wraps to 0
    }
    return make_uint4(u[0], u[1], u[2], __float_as_uint(q));
}

// squared minimum-image distance of two sites from their lattice coordinates
// (realcal.F90:271-285): the wrapped fractional difference to Cartesian
// (cell_normal is upper triangular), then for a skewed cell the reference's
// Z, Y, X shift sequence, which depends only on the difference modulo the
// lattice
__device__ __forceinline__ float lattice_d2(const FrameConst &fc, const uint4 &a, const uint4 &b)
{
    const float sx = (float)(int)(a.x - b.x), sy = (float)(int)(a.y - b.y), sz = (float)(int)(a.z - b.z);
    float dx, dy, dz;
    if (fc.is_cuboid) {
        dx = sx * fc.dsc[0];
        dy = sy * fc.dsc[4];
        dz = sz * fc.dsc[8];
    } else {
        dx = fmaf(fc.dsc[0], sx, fmaf(fc.dsc[3], sy, fc.dsc[6] * sz));
        dy = fmaf(fc.dsc[4], sy, fc.dsc[7] * sz);
        dz = fc.dsc[8] * sz;
        min_image(fc, dx, dy, dz);
    }
    return dx * dx + dy * dy + dz * dz;
}

// B-spline weights w[c] = M_ord(t + c), c = 0..ord-1, by the Cox-de Boor
// recursion M_k(x) = (x M_{k-1}(x) + (k-x) M_{k-1}(x-1)) / (k-1).  Same
// values as the explicit sum of spline.F90:46-58 without its cancellation.
__device__ __forceinline__ void bspline_weights(float t, int ord, float *w)
{
    w[0] = t;
    w[1] = 1.0f - t;
    for (int k = 3; k <= ord; ++k) {
        const float inv = 1.0f / (float)(k - 1);
        w[k - 1] = (1.0f - t) * w[k - 2] * inv;
        for (int c = k - 2; c >= 1; --c) w[c] = ((t + (float)c) * w[c] + ((float)(k - c) - t) * w[c - 1]) * inv;
        w[0] = t * w[0] * inv;
    }
}

// fractional grid coordinate of one site (recpcal.F90:517-536)
__device__ __forceinline__ void site_spline(const FrameConst &fc, float x, float y, float z, int K1, int K2,
                                            int K3, int ord, float *w /*[3*ord]*/, int4 &g)
{
    // the fraction and grid coordinate in FP64 from the FP32 position, so
    // that neither carries the rounding of a coordinate as large as the box
    // (or, for an unwrapped trajectory, several boxes); the weights in FP32
    const int Ks[3] = {K1, K2, K3};
    const double X = x, Y = y, Z = z;
    int gg[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double f = fc.invcld[k] * X + fc.invcld[k + 3] * Y + fc.invcld[k + 6] * Z;
        f -= floor(f);
        const double factor = f * (double)Ks[k];
        int rci = (int)factor;
        const float t = (float)(factor - (double)rci);
        if (rci >= Ks[k]) rci -= Ks[k];  // a fraction that rounded up to 1
        bspline_weights(t, ord, w + k * ord);
        gg[k] = rci;
    }
    g = make_int4(gg[0], gg[1], gg[2], 0);
}

// ---------------------------------------------------------------------------
// per-frame preparation
// ---------------------------------------------------------------------------

// realcal_prepare (realcal.F90:30-48) + recpcal_prepare_solvent
// (recpcal.F90:340-357) for every physical atom
// The frame's sort counters (cell counts, brick counts and the largest
// molecule reach) are cleared here too, before the kernels that add to them.
__global__ void k_prep_atoms(int natom, const float *__restrict__ pos, const float *__restrict__ charge,
                             FrameConst fc, int K1, int K2, int K3, int ord, float4 *__restrict__ posn,
                             uint4 *__restrict__ posu, float *__restrict__ splw, int4 *__restrict__ splg,
                             unsigned int *__restrict__ ccnt, int ncell, unsigned int *__restrict__ bcnt, int nbrick,
                             float *__restrict__ rmolmax)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int T = gridDim.x * blockDim.x;
    for (int i = a; i < ncell; i += T) ccnt[i] = 0u;
    for (int i = a; i < nbrick; i += T) bcnt[i] = 0u;
    if (a == 0) *rmolmax = 0.0f;
    if (a >= natom) return;
    const float x = pos[3 * (size_t)a], y = pos[3 * (size_t)a + 1], z = pos[3 * (size_t)a + 2];
    float wx = x, wy = y, wz = z;
    rotate_site(fc, wx, wy, wz);
    posn[a] = make_float4(wx, wy, wz, charge[a]);
    posu[a] = lattice_site(fc, x, y, z, charge[a]);
    float w[3 * ER_MAXORD];
    int4 g;
    site_spline(fc, x, y, z, K1, K2, K3, ord, w, g);
    for (int k = 0; k < 3 * ord; ++k) splw[(size_t)a * 3 * ord + k] = w[k];
    splg[a] = g;
}

// centre and reach radius of each molecule (for cutoff pruning)
#ifndef ER_SEG
#define ER_SEG 8  // atoms per molecule segment of the pair search's pruning
#endif

// centroid of the atoms [a0, a1) (minimum-image offsets from atom a0,
// averaged) and their largest minimum-image distance from it: a sphere that
// holds every atom, for the pair search's pruning (any centre is valid by
// the triangle inequality on minimum images; the centroid reaches least far
// for extended molecules)
__device__ __forceinline__ float4 atom_sphere(const FrameConst &fc, const float4 *__restrict__ posn, int a0, int a1)
{
    const float4 r = posn[a0];
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    for (int a = a0 + 1; a < a1; ++a) {
        const float4 p = posn[a];
        float dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
        min_image(fc, dx, dy, dz);
        ox += dx;
        oy += dy;
        oz += dz;
    }
    const float inv = 1.0f / (float)(a1 - a0);
    const float cx = r.x + ox * inv, cy = r.y + oy * inv, cz = r.z + oz * inv;
    float R = 0.0f;
    for (int a = a0; a < a1; ++a) {
        const float4 p = posn[a];
        float dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
        min_image(fc, dx, dy, dz);
        R = fmaxf(R, sqrtf(dx * dx + dy * dy + dz * dz));
    }
    return make_float4(cx, cy, cz, R);
}

// Per molecule: its bounding sphere (atom_sphere), and for every segment of
// ER_SEG consecutive atoms (molecule-relative) the segment's, stored at the
// segment's first atom in segref
__global__ void k_prep_mols(int nmol, const int *__restrict__ molbeg, const float4 *__restrict__ posn,
                            FrameConst fc, float4 *__restrict__ molref, float *__restrict__ rmolmax,
                            float4 *__restrict__ segref)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nmol) return;
    const int a0 = molbeg[j], a1 = molbeg[j + 1];
    const float4 ms = atom_sphere(fc, posn, a0, a1);
    const float R = ms.w;
    molref[j] = ms;
    for (int s0 = a0; s0 < a1; s0 += ER_SEG) {
        segref[s0] = atom_sphere(fc, posn, s0, min(s0 + ER_SEG, a1));
    }
    // frame maximum (non-negative floats order like their bit patterns)
    atomicMax(reinterpret_cast<unsigned int *>(rmolmax), __float_as_uint(R));
}

// Insertion (refs) or solute molecule (soln) sites: position, lattice
// coordinates and spline; k_prep_ins then adds the search sphere and the real-space self term
// realcal_self[_refs] (realcal.F90:669-771).  For soln the sites come from
// the frame.
__global__ void k_prep_sites(int nsite, int nslot, const SlotDesc *__restrict__ sd, const float *__restrict__ sltpos,
                             const float *__restrict__ frame_pos, const float *__restrict__ charge,
                             const int *__restrict__ molbeg, const int *__restrict__ tag, FrameConst fc, int K1,
                             int K2, int K3, int ord, float4 *__restrict__ insposn, uint4 *__restrict__ insposu,
                             float *__restrict__ insw, int4 *__restrict__ insg)
{
    // one thread per site of the run (every solute's insertions end to end):
    // spline weights, grid base, position and lattice coordinates.  The
    // first nslot threads also count the run for their solute (the run
    // number of the halt keys, read by the histograms later in the run).
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    if (is < nslot) *sd[is].seq += 1u;
    if (is >= nsite) return;
    const int k = slot_find(nslot, is, [&](int m) { return sd[m].site_off; });
    const int nslt = sd[k].nslt;
    const int local = is - sd[k].site_off;
    const int n = local / nslt, s = local % nslt;
    float x, y, z, q;
    if (tag) {  // soln: the solute molecules' sites from the frame
        const size_t a = (size_t)molbeg[tag[n]] + s;
        x = frame_pos[3 * a]; y = frame_pos[3 * a + 1]; z = frame_pos[3 * a + 2];
        q = charge[a];
    } else {  // refs: the run's insertion coordinates, solutes end to end
        const size_t o = (size_t)is * 3;
        x = sltpos[o]; y = sltpos[o + 1]; z = sltpos[o + 2];
        q = sd[k].sltq[s];
    }
    float w[3 * ER_MAXORD];
    int4 g;
    site_spline(fc, x, y, z, K1, K2, K3, ord, w, g);
    for (int c = 0; c < 3 * ord; ++c) insw[(size_t)is * 3 * ord + c] = w[c];
    insg[is] = g;
    insposu[is] = lattice_site(fc, x, y, z, q);
    rotate_site(fc, x, y, z);
    insposn[is] = make_float4(x, y, z, q);
}

// One workgroup of four waves per insertion (after k_prep_sites): search
// centre and reach radius (every wave computes them; wave 0 writes), the
// real-space self energy with the site pairs shared by the four waves, and
// the residual self term.
#define ER_PI_WAVES 4
__global__ void __launch_bounds__(64 * ER_PI_WAVES) k_prep_ins(int ntot, int nslot, const SlotDesc *__restrict__ sd,
                                                               FrameConst fc, float screen, int bare,
                                                               const float4 *__restrict__ insposn,
                                                               float4 *__restrict__ insref,
                                                               float *__restrict__ selfreal,
                                                               float *__restrict__ resself)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = blockIdx.x;
    if (n >= ntot) return;
    const int k = ER_SLOT_OF_INS(sd, nslot, n);
    const int local = n - sd[k].ins_off;
    if (local >= sd[k].nins) {  // padding of the solute's insertion range
        if (threadIdx.x == 0) {
            insref[n] = make_float4(0.f, 0.f, 0.f, 0.f);
            selfreal[n] = 0.0f;
            resself[n] = 0.0f;
        }
        return;
    }
    const int nslt = sd[k].nslt;
    const float4 *P = insposn + sd[k].site_off + (size_t)local * nslt;
    // The pair search's centre: the sites' centroid (their minimum-image
    // offsets from site 0, averaged), with reach R = the largest minimum-image
    // distance of a site from it.  Any centre bounds the search by the
    // triangle inequality on minimum images; the centroid of an extended
    // solute reaches far less than its first site (28 -> ~16 A for a
    // 150-site chain), and the search volume goes with the cube of cutoff +
    // reach.
    const float4 c0 = P[0];
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    for (int s = 1 + lane; s < nslt; s += 64) {
        float dx = P[s].x - c0.x, dy = P[s].y - c0.y, dz = P[s].z - c0.z;
        min_image(fc, dx, dy, dz);
        ox += dx;
        oy += dy;
        oz += dz;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ox += __shfl_xor(ox, off, 64);
        oy += __shfl_xor(oy, off, 64);
        oz += __shfl_xor(oz, off, 64);
    }
    const float inv = 1.0f / (float)nslt;
    const float cx = c0.x + ox * inv, cy = c0.y + oy * inv, cz = c0.z + oz * inv;
    float R = 0.0f;
    for (int s = lane; s < nslt; s += 64) {
        float dx = P[s].x - cx, dy = P[s].y - cy, dz = P[s].z - cz;
        min_image(fc, dx, dy, dz);
        R = fmaxf(R, sqrtf(dx * dx + dy * dy + dz * dz));
    }
    for (int off = 32; off > 0; off >>= 1) R = fmaxf(R, __shfl_xor(R, off, 64));
    if (bare) {  // cltype 0: realcal_self and residual_self_ene return at once
        if (threadIdx.x == 0) {
            insref[n] = make_float4(cx, cy, cz, R);
            selfreal[n] = 0.0f;
            resself[n] = 0.0f;
        }
        return;
    }
    // realcal_self_refs: -sum q^2 alpha/sqrt(pi) - sum_{i<j} q_i q_j erf(alpha r)/r,
    // in FP64 from the FP32 positions and charges: a 150-site solute has
    // 11 175 terms of up to tens of kcal/mol that largely cancel, and this
    // self energy enters the insertion's weight; computed once per
    // insertion, FP64 costs little here.  Wave w takes the rows s = w mod 4
    // and its lane l the pairs (s, t) with t - s - 1 = l mod 64, then a
    // fixed-order tree in each wave and the four waves added in order.
    double e = 0.0;
    const double rsqpi = 0.564189583547756286948079451560772586;  // 1/sqrt(pi)
    const double scr = screen;
    if (wave == 0)
        for (int s = lane; s < nslt; s += 64) e += -(double)P[s].w * (double)P[s].w * scr * rsqpi;
    for (int s = wave; s < nslt; s += ER_PI_WAVES) {
        const float4 a = P[s];
        for (int t = s + 1 + lane; t < nslt; t += 64) {
            const float4 b = P[t];
            double dx = (double)a.x - (double)b.x, dy = (double)a.y - (double)b.y, dz = (double)a.z - (double)b.z;
            min_image_d(fc, dx, dy, dz);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            e += -(double)a.w * (double)b.w * erf(scr * r) / r;
        }
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    __shared__ double wsum[ER_PI_WAVES];
    if (lane == 0) wsum[wave] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        e = wsum[0];
        for (int w = 1; w < ER_PI_WAVES; ++w) e += wsum[w];
        insref[n] = make_float4(cx, cy, cz, R);
        selfreal[n] = (float)e;
        // residual_self_ene (engproc.F90:967-977): -(pi Q^2 / alpha^2 / V) / 2 with
        // Q = mol_charge, a plain sum of the site charges (setconf.F90:838-840)
        float Q = 0.0f;
        for (int s = 0; s < nslt; ++s) Q += P[s].w;
        float epcl = ER_PI_F * Q * Q / screen / screen / fc.volume;
        epcl = epcl / 2.0f;
        resself[n] = 0.0f - epcl;
    }
}

// SPME influence function engfac (recpcal.F90:121-163), stored [m1][m3][m2]
__global__ void k_engfac(int K1, int K2, int K3, int ord, const double *__restrict__ s1,
                         const double *__restrict__ s2, const double *__restrict__ s3, double i00,
                         double i10, double i20, double i01, double i11, double i21, double i02, double i12,
                         double i22, double volume, double screen, float *__restrict__ engfac)
{
    const int H = K1 / 2 + 1;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)H * K2 * K3) return;
    const int m2 = idx % K2, m3 = (idx / K2) % K3, m1 = idx / ((size_t)K2 * K3);
    double f = 0.0;
    const int rc[3] = {m1, m2, m3}, Km[3] = {K1, K2, K3};
    bool skip = (m1 == 0 && m2 == 0 && m3 == 0);
    double inm[3];
    for (int m = 0; m < 3; ++m) {
        if ((ord % 2 == 1) && (2 * rc[m] == Km[m])) skip = true;
        inm[m] = (rc[m] <= Km[m] / 2) ? (double)rc[m] : (double)(rc[m] - Km[m]);
    }
    if (!skip) {
        // xs(m) = dot_product(invcl(:, m), inm)
        const double x0 = i00 * inm[0] + i10 * inm[1] + i20 * inm[2];
        const double x1 = i01 * inm[0] + i11 * inm[1] + i21 * inm[2];
        const double x2 = i02 * inm[0] + i12 * inm[1] + i22 * inm[2];
        const double r2 = x0 * x0 + x1 * x1 + x2 * x2;
        const double pi = 3.14159265358979323846;
        f = exp(-pi * pi * r2 / (screen * screen)) / r2 / pi / volume;
        f = f / (s1[m1] * s2[m2] * s3[m3]);
    }
    engfac[idx] = (float)f;
}

// PPPM optimal influence function (recpcal.F90:217-337), same layout as
// k_engfac.  For each wave vector k = 2 pi invcl^T m:
//   4 pi / V * sum_alias (k.km / |km|^2) exp(-|km|^2 / 4 a^2) U(km)^2
//   / (|k|^2 (s1 s2 s3)^2),  s_d = sum_l b_l sin^2(pi m_d / K_d)^l,
// aliases km = k + 2 pi invcl^T (mx K1, my K2, mz K2) over |m.| <= amax
// (the third alias component scales with K2, as the reference writes it).
// Computed in FP64, stored FP32; runs only when the cell changes.
struct PPPMCoef {
    double b[ER_MAXORD];  // gf_b(1:order)
    int order, a1, a2, a3;
};

__global__ void k_engfac_pppm(int K1, int K2, int K3, PPPMCoef pc, double i00, double i10, double i20, double i01,
                              double i11, double i21, double i02, double i12, double i22, double volume,
                              double screen, float *__restrict__ engfac)
{
    const int H = K1 / 2 + 1;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)H * K2 * K3) return;
    const int m2 = idx % K2, m3 = (idx / K2) % K3, m1 = idx / ((size_t)K2 * K3);
    const double pi = 3.14159265358979323846;
    const int rc[3] = {m1, m2, m3}, Km[3] = {K1, K2, K3};
    double inm[3], s2[3];
    for (int d = 0; d < 3; ++d) {
        inm[d] = (rc[d] <= Km[d] / 2) ? (double)rc[d] : (double)(rc[d] - Km[d]);
        const double sv = sin(pi * inm[d] / Km[d]);
        s2[d] = sv * sv;
    }
    const double k0 = 2.0 * pi * (i00 * inm[0] + i10 * inm[1] + i20 * inm[2]);
    const double k1 = 2.0 * pi * (i01 * inm[0] + i11 * inm[1] + i21 * inm[2]);
    const double k2c = 2.0 * pi * (i02 * inm[0] + i12 * inm[1] + i22 * inm[2]);
    const double kk = k0 * k0 + k1 * k1 + k2c * k2c;
    double f = 0.0;
    if (!(m1 == 0 && m2 == 0 && m3 == 0) && kk > 0.0) {
        auto u2p = [&](double t) {  // (sin t / t)^(2 order)
            if (t == 0.0) return 1.0;
            const double r = sin(t) / t, r2 = r * r;
            double u = 1.0;
            for (int l = 0; l < pc.order; ++l) u *= r2;
            return u;
        };
        const double inv4a2 = 0.25 / (screen * screen);
        double sum = 0.0;
        for (int mx = -pc.a1; mx <= pc.a1; ++mx) {
            const double u1 = u2p(pi * (inm[0] / K1 + mx));
            for (int my = -pc.a2; my <= pc.a2; ++my) {
                const double u2 = u2p(pi * (inm[1] / K2 + my));
                for (int mz = -pc.a3; mz <= pc.a3; ++mz) {
                    const double u3 = u2p(pi * (inm[2] / K3 + mz));
                    const double n0 = (double)mx * K1, n1 = (double)my * K2, n2 = (double)mz * K2;
                    const double q0 = k0 + 2.0 * pi * (i00 * n0 + i10 * n1 + i20 * n2);
                    const double q1 = k1 + 2.0 * pi * (i01 * n0 + i11 * n1 + i21 * n2);
                    const double q2 = k2c + 2.0 * pi * (i02 * n0 + i12 * n1 + i22 * n2);
                    const double g = exp(-q0 * q0 * inv4a2) * exp(-q1 * q1 * inv4a2) * exp(-q2 * q2 * inv4a2);
                    sum += (k0 * q0 + k1 * q1 + k2c * q2) * g * (u1 * u2 * u3) / (q0 * q0 + q1 * q1 + q2 * q2);
                }
            }
        }
        double s1 = 0.0, sy = 0.0, sz = 0.0;
        for (int l = pc.order - 1; l >= 0; --l) {
            s1 = pc.b[l] + s1 * s2[0];
            sy = pc.b[l] + sy * s2[1];
            sz = pc.b[l] + sz * s2[2];
        }
        const double den = (s1 * sy * sz) * (s1 * sy * sz);
        f = sum * 4.0 * pi / (kk * den) / volume;
    }
    engfac[idx] = (float)f;
}

// ---------------------------------------------------------------------------
// PME, stage A: per (insertion, m1) plane.  The solute's structure factor is
// generated directly from its separable B-spline factors,
//   rho^(m) = sum_s q_s F1_s(m1) F2_s(m2) F3_s(m3),
//   Fd_s(m) = sum_c W_s[d][c] exp(-2 pi i m (g_d - c)/K_d),
// which equals the R2C transform of the spread grid (recpcal.F90:453-475).
// The FP64 half-plane self energy (recpcal.F90:482-491) and the engfac
// multiply (recpcal.F90:493-497) are fused, then the plane is inverse
// transformed along axis 3 and axis 2 in LDS by a two-pass (four-step)
// split N = R1*R2 of each line: pass A runs R2-strided R1-point DFTs and the
// W_N^{n2 k1} twiddle, pass B contiguous R2-point DFTs; every pass keeps all
// 256 threads busy on small register DFTs.
// T layout: [chunk][H][K2][K3] (x3 fastest).
// ---------------------------------------------------------------------------
#define ER_SCH 8  // solute sites per LDS table chunk

// offsets of the small-radix tables exp(+2 pi i k/r), r = 4..10, concatenated
__host__ __device__ constexpr int tws_off(int r) { return r * (r - 1) / 2 - 6; }
#define ER_TWS_TOTAL 130  // radices 4..16

// Separable structure-factor factors of every insertion site, computed once
// per run instead of once per (insertion, m1) plane:
//   F_axis(m) = sum_c w_c exp(-2 pi i m (g - c) / K)   (axis 1 scaled by q)
// Ftab: [insertion][site][H + K2 + K3]: m1 = 0..H-1, then m2, then m3.
__global__ void k_ins_factors(int nrows, int ord, int K1, int K2, int K3, int H,
                              const float4 *__restrict__ insposn, const float *__restrict__ insw,
                              const int4 *__restrict__ insg, const float2 *__restrict__ tw1,
                              const float2 *__restrict__ tw2, const float2 *__restrict__ tw3,
                              float2 *__restrict__ Ftab)
{
    const int per = H + K2 + K3;
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (size_t)nrows * per) return;
    const int r = (int)(f % per);
    const size_t site = f / per;
    const int4 g = insg[site];
    int axis, m, K, gd;
    const float2 *tw;
    if (r < H) { axis = 0; m = r; K = K1; gd = g.x; tw = tw1; }
    else if (r < H + K2) { axis = 1; m = r - H; K = K2; gd = g.y; tw = tw2; }
    else { axis = 2; m = r - H - K2; K = K3; gd = g.z; tw = tw3; }
    const float *w = insw + site * 3 * ord + axis * ord;
    float2 a = make_float2(0.f, 0.f);
    // e = m (g - c) mod K: one reduction for c = 0 (m, pmod(.) < K <= 512, so
    // the product fits 32 bits), then e - m mod K for each next c
    int e = (int)(((unsigned int)m * (unsigned int)pmod(gd, K)) % (unsigned int)K);
    for (int c = 0; c < ord; ++c) {
        const float2 t = tw[e];  // exp(+2 pi i e/K); the forward sum uses the conjugate
        a.x += w[c] * t.x;
        a.y -= w[c] * t.y;
        e -= m;
        e += (e < 0) ? K : 0;
    }
    if (axis == 0) a = erfft::cscale(a, insposn[site].w);
    Ftab[f] = a;
}

// One term engfac(m) |rho^(m)|^2 of the half-plane self energy
// (recpcal.F90:482-491): formed in FP32 (every term is non-negative, so
// the plane's sum carries at most a few FP32 roundings of relative error)
// and accumulated in FP64 over the plane.
__device__ __forceinline__ double self_term(float f, float2 a) { return (double)(f * fmaf(a.x, a.x, a.y * a.y)); }
// weight of plane m1 of the half spectrum: 1/2 for the planes that are their
// own mirror image (m1 = 0 and, for even K1, m1 = K1/2), 1 otherwise
__device__ __forceinline__ double plane_weight(int K1, int m1)
{
    return (m1 == 0 || (K1 % 2 == 0 && m1 == K1 / 2)) ? 0.5 : 1.0;
}

// In-place 2-D DFT of the LDS plane [K2][PS] (x3/m3 fastest) along axis 3
// (K3 = R1*R2) and then axis 2 (K2 = C1*C2), each as a four-step transform:
// pass A = R1-point DFTs over stride-R2 elements + twiddles, pass B = R2-point
// DFTs with the results written in natural order.  SIGN = +1 is the inverse
// (e^{+i}) transform, -1 the forward one; both unnormalised.
template <int K2, int K3, int R1, int R2, int C1, int C2, int NT, int SIGN>
__device__ __forceinline__ void plane_dft(float2 *plane, const float2 *t2, const float2 *t3,
                                          const float2 *__restrict__ tws, int tid,
                                          float2 *__restrict__ gout = nullptr)
{
    constexpr int PS = K3 + 1;
    // the pass-A twiddle index n2 * k1 (n2 < R2, k1 < R1) stays below
    // (R1 - 1) (R2 - 1) < K3, so it needs no reduction modulo the length
    static_assert((R1 - 1) * (R2 - 1) < K3 && (C1 - 1) * (C2 - 1) < K2, "twiddle index range");
    // ---- rows (axis 3), pass A in place: each thread's items are loaded
    // together, then transformed, then stored (their LDS latencies overlap).
    // Items run rows fastest, so a wave's lanes share n2 and read one
    // twiddle t3[n2 k1] ----
    {
        constexpr int IA = (K2 * R2 + NT - 1) / NT;
        float2 v[IA][R1];
#pragma unroll
        for (int j = 0; j < IA; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < K2 * R2) {
                const int m2 = it % K2, n2 = it / K2;
#pragma unroll
                for (int a = 0; a < R1; ++a) v[j][a] = plane[m2 * PS + R2 * a + n2];
            }
        }
#pragma unroll
        for (int j = 0; j < IA; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < K2 * R2) {
                const int m2 = it % K2, n2 = it / K2;
                erfft::fft<R1, SIGN>(v[j], tws + tws_off(R1));
#pragma unroll
                for (int k1 = 1; k1 < R1; ++k1) v[j][k1] = erfft::twmul<SIGN>(v[j][k1], t3[n2 * k1]);
#pragma unroll
                for (int k1 = 0; k1 < R1; ++k1) plane[m2 * PS + R2 * k1 + n2] = v[j][k1];
            }
        }
    }
    __syncthreads();
    // ---- rows, pass B (outputs held until every thread has read) ----
    {
        constexpr int IB = (K2 * R1 + NT - 1) / NT;
        float2 hold[IB][R2];
#pragma unroll
        for (int j = 0; j < IB; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < K2 * R1) {
                const int m2 = it % K2, k1 = it / K2;  // rows fastest: lanes one row apart, no bank conflicts
#pragma unroll
                for (int b = 0; b < R2; ++b) hold[j][b] = plane[m2 * PS + R2 * k1 + b];
                erfft::fft<R2, SIGN>(hold[j], tws + tws_off(R2));
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IB; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < K2 * R1) {
                const int m2 = it % K2, k1 = it / K2;
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) plane[m2 * PS + k1 + R1 * k2] = hold[j][k2];
            }
        }
    }
    __syncthreads();
    // ---- columns (axis 2), pass A in place ----
    {
        constexpr int IA = (C2 * K3 + NT - 1) / NT;
        float2 v[IA][C1];
#pragma unroll
        for (int j = 0; j < IA; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < C2 * K3) {
                const int n2 = it / K3, x3 = it % K3;
#pragma unroll
                for (int a = 0; a < C1; ++a) v[j][a] = plane[(C2 * a + n2) * PS + x3];
            }
        }
#pragma unroll
        for (int j = 0; j < IA; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < C2 * K3) {
                const int n2 = it / K3, x3 = it % K3;
                erfft::fft<C1, SIGN>(v[j], tws + tws_off(C1));
#pragma unroll
                for (int k1 = 1; k1 < C1; ++k1) v[j][k1] = erfft::twmul<SIGN>(v[j][k1], t2[n2 * k1]);
#pragma unroll
                for (int k1 = 0; k1 < C1; ++k1) plane[(C2 * k1 + n2) * PS + x3] = v[j][k1];
            }
        }
    }
    __syncthreads();
    // ---- columns, pass B ----
    {
        constexpr int IC = (C1 * K3 + NT - 1) / NT;
        float2 hold[IC][C2];
#pragma unroll
        for (int j = 0; j < IC; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < C1 * K3) {
                const int k1 = it / K3, x3 = it % K3;
#pragma unroll
                for (int b = 0; b < C2; ++b) hold[j][b] = plane[(C2 * k1 + b) * PS + x3];
                erfft::fft<C2, SIGN>(hold[j], tws + tws_off(C2));
            }
        }
        if (gout) {
            // results straight to global memory as [x2][x3] (x3 fastest:
            // consecutive lanes write consecutive x3)
#pragma unroll
            for (int j = 0; j < IC; ++j) {
                const unsigned int it = (unsigned int)tid + j * NT;
                if (it < C1 * K3) {
                    const int k1 = it / K3, x3 = it % K3;
#pragma unroll
                    for (int k2 = 0; k2 < C2; ++k2) gout[(size_t)(k1 + C1 * k2) * K3 + x3] = hold[j][k2];
                }
            }
            return;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IC; ++j) {
            const unsigned int it = (unsigned int)tid + j * NT;
            if (it < C1 * K3) {
                const int k1 = it / K3, x3 = it % K3;
#pragma unroll
                for (int k2 = 0; k2 < C2; ++k2) plane[(k1 + C1 * k2) * PS + x3] = hold[j][k2];
            }
        }
    }
    __syncthreads();
}

// PME, stage A: one (insertion, m1) plane of the half spectrum per workgroup.
//   1. rho^(m1, m2, m3) = FFT of the solute's spread charges.  Few sites:
//      sum_s F1_s(m1) F2_s(m2) F3_s(m3) from the separable factor table
//      (nslt * K2 * K3 complex products).  Many sites (sparse = 1): the
//      (x2, x3) plane of sum_s F1_s(m1) w2_s(x2) w3_s(x3) is scattered in
//      LDS (ORD^2 points per site, sites in order) and transformed forward
//      along axes 3 and 2 (the same transform as the spread + forward FFT
//      of recpcal.F90:446-478, done per plane).
//   2. FP64 half-plane self-energy partial and the engfac multiply.
//   3. inverse DFT along axes 3 and 2; T[nl][m1][x2][x3] for stage B.
template <int K2, int K3, int R1, int R2, int C1, int C2, int NT, bool SPARSE>
__global__ void __launch_bounds__(NT) k_pme_spec2d(int nslot, const SlotDesc *__restrict__ sd, int ord, int K1,
                                                    int n_off,
                                                    const float2 *__restrict__ Ftab,
                                                    const float *__restrict__ insw, const int4 *__restrict__ insg,
                                                    const float *__restrict__ engfac,
                                                    const float2 *__restrict__ tw2, const float2 *__restrict__ tw3,
                                                    const float2 *__restrict__ tws, float2 *__restrict__ T,
                                                    double *__restrict__ selfpart)
{
    static_assert(R1 * R2 == K3 && C1 * C2 == K2, "line splits");
    constexpr int PS = K3 + 1;  // padded plane row stride
    constexpr int NE = K2 * K3;
    constexpr int EPT = (NE + NT - 1) / NT;
    extern __shared__ __align__(16) unsigned char smem[];
    float2 *plane = reinterpret_cast<float2 *>(smem);  // [K2][PS]
    float2 *F2 = plane + K2 * PS;                       // [SCH][K2]: F1(m1) * F2(m2)
    float2 *F3 = F2 + ER_SCH * K2;                      // [SCH][K3]
    float2 *t2 = F3 + ER_SCH * K3;                      // [K2]
    float2 *t3 = t2 + K2;                               // [K3]
    double *red = reinterpret_cast<double *>(t3 + K3);  // [NT/64]

    const int H = K1 / 2 + 1;
    const int m1 = blockIdx.x;
    const int nl = blockIdx.y;
    const int n = n_off + nl;
    const int tid = threadIdx.x;
    const int ks = ER_SLOT_OF_INS(sd, nslot, n);
    const int local = n - sd[ks].ins_off;
    if (local >= sd[ks].nins) {
        // padding insertion: a zero plane and no self energy
        float2 *Tp = T + ((size_t)nl * H + m1) * (size_t)NE;
        for (int e = tid; e < NE; e += NT) Tp[e] = make_float2(0.f, 0.f);
        if (tid == 0) selfpart[(size_t)n * H + m1] = 0.0;
        return;
    }
    const int nslt = sd[ks].nslt;
    const size_t rowb = (size_t)sd[ks].site_off + (size_t)local * nslt;  // the insertion's first site row
    for (int k = tid; k < K2; k += NT) t2[k] = tw2[k];
    for (int k = tid; k < K3; k += NT) t3[k] = tw3[k];
    const int per = H + K2 + K3;
    double sp = 0.0;  // this thread's share of the FP64 half-plane self energy

    if constexpr (!SPARSE) {
        // ---- separable generation, sites in LDS chunks: thread t owns row
        // m2 = t % K2 and the m3 = g, g + GR, ... of group g = t / K2, so the
        // row factor is read once per site and the column factors are shared
        // by the wave's lanes ----
        constexpr int GR = NT / K2;                  // thread groups per plane (>= 1)
        constexpr int EPG = (K3 + GR - 1) / GR;      // elements per thread
        const int m2 = tid % K2, g = tid / K2;
        const bool owner = g < GR;
        float2 acc[EPG];
#pragma unroll
        for (int k = 0; k < EPG; ++k) acc[k] = make_float2(0.f, 0.f);
        for (int s0 = 0; s0 < nslt; s0 += ER_SCH) {
            const int ns = min(ER_SCH, nslt - s0);
            for (int f = tid; f < ns * (K2 + K3); f += NT) {
                const int s = f / (K2 + K3), r = f % (K2 + K3);
                const float2 *Fs = Ftab + (rowb + s0 + s) * per;
                if (r < K2) F2[s * K2 + r] = erfft::cmul(Fs[m1], Fs[H + r]);
                else F3[s * K3 + (r - K2)] = Fs[H + r];
            }
            __syncthreads();
            if (owner) {
                for (int s = 0; s < ns; ++s) {
                    const float2 B = F2[s * K2 + m2];
#pragma unroll
                    for (int k = 0; k < EPG; ++k) {
                        const int m3 = g + k * GR;
                        if (m3 < K3) {
                            const float2 f3 = F3[s * K3 + m3];
                            // four chained FMAs per complex product
                            acc[k].x = fmaf(B.x, f3.x, acc[k].x);
                            acc[k].x = fmaf(-B.y, f3.y, acc[k].x);
                            acc[k].y = fmaf(B.x, f3.y, acc[k].y);
                            acc[k].y = fmaf(B.y, f3.x, acc[k].y);
                        }
                    }
                }
            }
            __syncthreads();
        }
        // engfac multiply and half-plane self energy on the registers, then
        // one store of the plane
        const float *ef = engfac + (size_t)m1 * NE;
        if (owner) {
#pragma unroll
            for (int k = 0; k < EPG; ++k) {
                const int m3 = g + k * GR;
                if (m3 < K3) {
                    const float f = ef[m3 * K2 + m2];  // [m1][m3][m2]
                    const float2 a = acc[k];
                    sp += self_term(f, a);
                    plane[m2 * PS + m3] = erfft::cscale(a, f);
                }
            }
        }
    } else {
        // ---- sparse spread of the plane, then forward DFT along axes 3, 2 ----
        for (int e = tid; e < K2 * PS; e += NT) plane[e] = make_float2(0.f, 0.f);
        // the sites' factors F1(m1), grid bases and axis-2/3 weights are
        // staged in LDS (in the factor-table area, unused on this path) by the
        // whole workgroup, chunk by chunk, so that the scatter below reads no
        // global memory
        float2 *sA = F2;
        const int CH = (int)((size_t)ER_SCH * (K2 + K3) * sizeof(float2) / (sizeof(float2) + sizeof(int2) +
                                                                             2 * ord * sizeof(float)));
        int2 *sg = reinterpret_cast<int2 *>(sA + CH);
        float *sw = reinterpret_cast<float *>(sg + CH);  // [CH][2 ord]: w2 then w3
        for (int s0 = 0; s0 < nslt; s0 += CH) {
            const int ns = min(CH, nslt - s0);
            __syncthreads();
            for (int f = tid; f < ns; f += NT) {
                const size_t site = rowb + s0 + f;
                const int4 g = insg[site];
                sA[f] = Ftab[site * per + m1];
                sg[f] = make_int2(g.y, g.z);
            }
            for (int f = tid; f < ns * 2 * ord; f += NT) {
                const int sl = f / (2 * ord), c = f % (2 * ord);
                sw[f] = insw[(rowb + s0 + sl) * 3 * ord + ord + c];
            }
            __syncthreads();
            // one site's stencil point (c2, c3) into the plane (the grid
            // bases lie in [0, K), so one conditional add wraps an index)
            auto add_point = [&](int s, int c2, int c3) {
                const int2 g = sg[s];
                const float2 A = sA[s];
                int x2 = g.x - c2, x3 = g.y - c3;
                x2 += (x2 < 0) ? K2 : 0;
                x3 += (x3 < 0) ? K3 : 0;
                const float2 v = erfft::cscale(A, sw[s * 2 * ord + c2] * sw[s * 2 * ord + ord + c3]);
                float2 *q = plane + x2 * PS + x3;
                *q = make_float2(q->x + v.x, q->y + v.y);
            };
            const int Q = (ns + 3) >> 2;
            if (ord * ord <= 16 && (Q + 31) / 32 <= NT / 64) {
                // one wave, four sites at a time: lane (k, p) adds point p of
                // site j + k Q (Q = a quarter of the chunk, so that bonded
                // neighbours fall in different groups).  The ORD^2 points of a
                // site are distinct; when no two of the group's ORD x ORD
                // blocks meet (circular distance of their corners below ORD
                // on both axes) the four sites go in one step, else one after
                // another: in two steps, k = 0, 2 then k = 1, 3, when neither
                // pair of a step meets, else in k order.  Every point
                // therefore receives its additions in an order fixed by the
                // sites alone.  The groups' flags are found first by the whole
                // workgroup, one thread per (group, pair of sites), as bits in
                // the self-energy reduction array (unused until the end):
                // meet = some pair meets, far = pair (0, 2) or (1, 3) meets.
                const int nw = (Q + 31) / 32;                                  // words per bit set
                unsigned int *meet = reinterpret_cast<unsigned int *>(red);  // [nw], then far [nw]
                unsigned int *far = meet + nw;
                if (tid < 2 * nw) meet[tid] = 0u;
                __syncthreads();
                for (int t = tid; t < Q * 6; t += NT) {
                    const int jg = t / 6, pr = t % 6;
                    const int u = pr < 3 ? 0 : (pr < 5 ? 1 : 2);
                    const int v = pr < 3 ? pr + 1 : (pr < 5 ? pr - 1 : 3);
                    const int su = jg + u * Q, sv = jg + v * Q;
                    if (su < ns && sv < ns) {
                        const int2 a = sg[su], b = sg[sv];
                        int dx = a.x - b.x, dy = a.y - b.y;
                        dx += (dx < 0) ? K2 : 0;
                        dy += (dy < 0) ? K3 : 0;
                        dx = min(dx, K2 - dx);
                        dy = min(dy, K3 - dy);
                        if (dx < ord && dy < ord) {
                            atomicOr(&meet[jg >> 5], 1u << (jg & 31));
                            if (v == u + 2) atomicOr(&far[jg >> 5], 1u << (jg & 31));
                        }
                    }
                }
                __syncthreads();
                if (tid < 64) {
                    const int k = tid >> 4, pnt = tid & 15;
                    const bool on = pnt < ord * ord;
                    const int c2 = on ? pnt / ord : 0, c3 = on ? pnt % ord : 0;
                    for (int jg = 0; jg < Q; ++jg) {
                        const bool clash = (meet[jg >> 5] >> (jg & 31)) & 1u;
                        const bool clash_far = (far[jg >> 5] >> (jg & 31)) & 1u;
                        const int s = jg + k * Q;
                        const bool mine = on && s < ns;
                        // lanes of different k may add to one point in
                        // consecutive steps: a wavefront fence between steps
                        // keeps the compiler from merging or reordering them
                        // (the wave executes them in order)
                        if (!clash) {
                            if (mine) add_point(s, c2, c3);
                        } else if (!clash_far) {
                            if (mine && (k & 1) == 0) add_point(s, c2, c3);
                            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                            if (mine && (k & 1) == 1) add_point(s, c2, c3);
                        } else {
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                if (k == u && mine) add_point(s, c2, c3);
                                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                            }
                        }
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    }
                }
            } else if (tid < ord * ord) {
                // larger orders (or chunks past the flag words): one lane per
                // stencil point, sites in order
                const int c2 = tid / ord, c3 = tid % ord;
                for (int s = 0; s < ns; ++s) add_point(s, c2, c3);
            }
        }
        __syncthreads();
        plane_dft<K2, K3, R1, R2, C1, C2, NT, -1>(plane, t2, t3, tws, tid);
    }
    if constexpr (SPARSE) {
        __syncthreads();
        // ---- engfac multiply + half-plane self energy ----
        const float *ef = engfac + (size_t)m1 * NE;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = tid + k * NT;
            if (e < NE) {
                const int m2 = e % K2, m3 = e / K2;
                const float f = ef[e];  // [m1][m3][m2]
                float2 *q = plane + m2 * PS + m3;
                const float2 a = *q;
                sp += self_term(f, a);
                *q = erfft::cscale(a, f);
            }
        }
    }
    __syncthreads();
    // ---- inverse DFT along axes 3 and 2, written to T[nl][m1][x2][x3] ----
    plane_dft<K2, K3, R1, R2, C1, C2, NT, +1>(plane, t2, t3, tws, tid, T + ((size_t)nl * H + m1) * (size_t)NE);

    for (int off = 32; off > 0; off >>= 1) sp += __shfl_down(sp, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = sp;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int wv = 0; wv < NT / 64; ++wv) t += red[wv];
        selfpart[(size_t)n * H + m1] = plane_weight(K1, m1) * t;
    }
}

// ---------------------------------------------------------------------------
// PME, stage B: complex-to-real inverse DFT along axis 1 of NI insertions x
// R lines (x2 range) of one x3, via a half-length complex transform:
//   Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/K1} (X[k] - conj X[M-k]),
//   z = IDFT_M(Z),  phi[2j] = Re z[j],  phi[2j+1] = Im z[j],  M = K1/2.
// One line per thread (insertion fastest across lanes), so the output is
// written directly as 64-byte runs of NI interleaved insertions.
// ---------------------------------------------------------------------------
template <int K1, int NI>
__global__ void __launch_bounds__(256) k_pme_axis1(int nchunk, int K2, int K3, int R,
                                                   const float2 *__restrict__ T, float *__restrict__ phi,
                                                   const float2 *__restrict__ tw1, const float2 *__restrict__ twm)
{
    constexpr int M = K1 / 2;
    constexpr int H = M + 1;
    // lines (x2 = blockIdx.y, x3 = x3_0 + r): consecutive x3 are adjacent in T
    const int tile = blockIdx.z, x2 = blockIdx.y, x3_0 = blockIdx.x * R;
    const int tid = threadIdx.x;
    const int i = tid % NI, r = tid / NI;
    if (r >= R) return;
    const int x3 = x3_0 + r;
    const int nl = tile * NI + i;
    const size_t plane = (size_t)K2 * K3;
    float2 z[M];
    if (nl < nchunk) {
        const float2 *src = T + (size_t)nl * H * plane + (size_t)x2 * K3 + x3;
        float2 X[H];
#pragma unroll
        for (int m = 0; m < H; ++m) X[m] = src[(size_t)m * plane];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const float2 a = X[k], b = X[M - k];
            const float2 e = make_float2(a.x + b.x, a.y - b.y);            // X[k] + conj X[M-k]
            const float2 d = make_float2(a.x - b.x, a.y + b.y);            // X[k] - conj X[M-k]
            const float2 o = (k == 0) ? d : erfft::cmul(d, tw1[k]);
            z[k] = make_float2(e.x - o.y, e.y + o.x);                     // e + i o
        }
        erfft::fft<M, +1>(z, twm);
    } else {
#pragma unroll
        for (int k = 0; k < M; ++k) z[k] = make_float2(0.f, 0.f);
    }
    const size_t G = plane * K1;
    float *dst = phi + (size_t)tile * G * NI + (((size_t)x2 * K3 + x3) * K1) * NI + i;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        dst[(2 * j) * NI] = z[j].x;
        dst[(2 * j + 1) * NI] = z[j].y;
    }
}

// Batched form of stage B: RL lines (consecutive x3, one x2) x 16
// insertions per workgroup.  Lanes run along x3 first, so a wave reads
// RL * 8-byte runs of T; the lines are transposed through LDS ([line][x1][16],
// rows padded by 8 floats: the transposed writes hit distinct banks) and each
// leaves as one contiguous K1 * 16-float run of phi written with 16-byte
// stores.  Per-line arithmetic as in k_pme_axis1.  XCD-aware order: workgroup
// b runs on XCD b % 8, and each XCD sweeps one contiguous eighth of the
// (x3 group, x2, tile) blocks with x3 groups fastest, so the groups sharing a
// 128-byte line of T are fetched back to back through the same L2.
template <int K1, int RL>
__global__ void __launch_bounds__(RL * 16) k_pme_axis1_t(int nchunk, int K2, int K3,
                                                          const float2 *__restrict__ T, float *__restrict__ phi,
                                                          const float2 *__restrict__ tw1,
                                                          const float2 *__restrict__ twm)
{
    constexpr int NI = 16, M = K1 / 2, H = M + 1, LS = K1 * NI + 8;
    extern __shared__ __align__(16) float sml[];  // [RL][LS]
    const int ngx = K3 / RL;
    const int total = ngx * K2 * ((nchunk + NI - 1) / NI);
    const int per = (total + 7) / 8;
    const int blk = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (blk >= total) return;
    const int tile = blk / (ngx * K2), x2 = (blk / ngx) % K2, x3_0 = (blk % ngx) * RL;
    const int tid = threadIdx.x;
    const int r = tid % RL, i = tid / RL;
    const int nl = tile * NI + i;
    const size_t plane = (size_t)K2 * K3;
    float2 z[M];
    if (nl < nchunk) {
        const float2 *src = T + (size_t)nl * H * plane + (size_t)x2 * K3 + x3_0 + r;
        float2 X[H];
#pragma unroll
        for (int m = 0; m < H; ++m) X[m] = src[(size_t)m * plane];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const float2 a = X[k], b = X[M - k];
            const float2 e = make_float2(a.x + b.x, a.y - b.y);            // X[k] + conj X[M-k]
            const float2 d = make_float2(a.x - b.x, a.y + b.y);            // X[k] - conj X[M-k]
            const float2 o = (k == 0) ? d : erfft::cmul(d, tw1[k]);
            z[k] = make_float2(e.x - o.y, e.y + o.x);                     // e + i o
        }
        erfft::fft<M, +1>(z, twm);
    } else {
#pragma unroll
        for (int k = 0; k < M; ++k) z[k] = make_float2(0.f, 0.f);
    }
    float *o = sml + r * LS + i;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        o[(2 * j) * NI] = z[j].x;
        o[(2 * j + 1) * NI] = z[j].y;
    }
    __syncthreads();
    const size_t G = plane * K1;
    constexpr int PERLINE = K1 * NI / 4;
    // the RL lines (x2, x3_0 .. x3_0 + RL - 1) are one contiguous block of
    // phi [x2][x3][x1][16]
    float4 *dst = reinterpret_cast<float4 *>(phi + (size_t)tile * G * NI) + ((size_t)x2 * K3 + x3_0) * PERLINE;
    for (int f = tid; f < RL * PERLINE; f += RL * 16) {
        const int line = f / PERLINE, rem = f % PERLINE;
        dst[f] = *reinterpret_cast<const float4 *>(sml + line * LS + rem * 4);
    }
}

// ---------------------------------------------------------------------------
// PME for grid shapes without a compiled plane or line kernel (any K, any
// factors): the same two stages with direct DFTs, each output a sum over its
// line with twiddles e^{+2 pi i k / K} from the tables, O(K) work per
// element instead of O(log K).  The reference transforms any K1 x K2 x K3
// (erdst/fft_iface.F90:38-50); these kernels keep that generality.
// ---------------------------------------------------------------------------
#define ER_GEN_NT 256

// nline lines of length K in LDS (line L at base + L * lstride, element m at
// + m * estride), transformed in place: out(x) = sum_m v(m) e^{SIGN 2 pi i m x / K}.
// One wave per line; lane l computes the outputs x = l, l + 64, ... (K <= 256)
// in registers before they are stored, so a line is read whole first.
template <int SIGN>
__device__ __forceinline__ void lds_dft_lines(float2 *base, int nline, int lstride, int estride, int K,
                                              const float2 *tw, int tid)
{
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int NW = ER_GEN_NT / 64;
    for (int L = wave; L < nline; L += NW) {
        float2 *line = base + (size_t)L * lstride;
        float2 out[4];
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int x = lane + 64 * o;
            float2 acc = make_float2(0.f, 0.f);
            if (x < K) {
                int idx = 0;  // (m * x) mod K
                for (int m = 0; m < K; ++m) {
                    const float2 v = line[m * estride];
                    const float2 t = tw[idx];
                    const float ty = SIGN > 0 ? t.y : -t.y;
                    acc.x += v.x * t.x - v.y * ty;
                    acc.y += v.x * ty + v.y * t.x;
                    idx += x;
                    if (idx >= K) idx -= K;
                }
            }
            out[o] = acc;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int x = lane + 64 * o;
            if (x < K) line[x * estride] = out[o];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// LDS bytes of k_pme_spec2d_gen for a K2 x K3 plane
__host__ __device__ constexpr size_t spec2d_gen_lds(int K2, int K3)
{
    return ((size_t)K2 * (K3 + 1) + K2 + K3 + (size_t)ER_SCH * (K2 + K3)) * 8 + (ER_GEN_NT / 64) * 8;
}

// Stage A for any K2 x K3 (one (insertion, m1) plane per workgroup): the
// structure factor (separable product, or for many-site solutes the scatter
// + forward DFT along axes 3 and 2), the FP64 half-plane self energy and the
// engfac multiply, the inverse DFT along axes 3 and 2, T[nl][m1][x2][x3].
__global__ void __launch_bounds__(ER_GEN_NT) k_pme_spec2d_gen(int nslot, const SlotDesc *__restrict__ sd, int ord,
                                                              int K1, int K2, int K3, int sparse, int n_off,
                                                              const float2 *__restrict__ Ftab,
                                                              const float *__restrict__ insw,
                                                              const int4 *__restrict__ insg,
                                                              const float *__restrict__ engfac,
                                                              const float2 *__restrict__ tw2,
                                                              const float2 *__restrict__ tw3, float2 *__restrict__ T,
                                                              double *__restrict__ selfpart)
{
    constexpr int NT = ER_GEN_NT;
    extern __shared__ __align__(16) unsigned char smem[];
    const int PS = K3 + 1, NE = K2 * K3, H = K1 / 2 + 1;
    float2 *plane = reinterpret_cast<float2 *>(smem);  // [K2][PS]
    float2 *t2 = plane + (size_t)K2 * PS;               // [K2]
    float2 *t3 = t2 + K2;                               // [K3]
    float2 *F2 = t3 + K3;                               // [SCH][K2]: F1(m1) F2(m2)
    float2 *F3 = F2 + ER_SCH * K2;                      // [SCH][K3]
    double *red = reinterpret_cast<double *>(F3 + ER_SCH * K3);
    const int m1 = blockIdx.x, nl = blockIdx.y, n = n_off + nl, tid = threadIdx.x;
    float2 *Tp = T + ((size_t)nl * H + m1) * (size_t)NE;
    const int ks = ER_SLOT_OF_INS(sd, nslot, n);
    const int local = n - sd[ks].ins_off;
    if (local >= sd[ks].nins) {  // padding insertion: a zero plane and no self energy
        for (int e = tid; e < NE; e += NT) Tp[e] = make_float2(0.f, 0.f);
        if (tid == 0) selfpart[(size_t)n * H + m1] = 0.0;
        return;
    }
    const int nslt = sd[ks].nslt;
    const size_t rowb = (size_t)sd[ks].site_off + (size_t)local * nslt;
    const int per = H + K2 + K3;
    for (int k = tid; k < K2; k += NT) t2[k] = tw2[k];
    for (int k = tid; k < K3; k += NT) t3[k] = tw3[k];
    for (int e = tid; e < K2 * PS; e += NT) plane[e] = make_float2(0.f, 0.f);
    __syncthreads();
    if (!sparse) {
        // element e = m2 * K3 + m3 owned by one thread, sites added in order
        for (int s0 = 0; s0 < nslt; s0 += ER_SCH) {
            const int ns = min(ER_SCH, nslt - s0);
            for (int f = tid; f < ns * (K2 + K3); f += NT) {
                const int s = f / (K2 + K3), r = f % (K2 + K3);
                const float2 *Fs = Ftab + (rowb + s0 + s) * per;
                if (r < K2) F2[s * K2 + r] = erfft::cmul(Fs[m1], Fs[H + r]);
                else F3[s * K3 + (r - K2)] = Fs[H + r];
            }
            __syncthreads();
            for (int e = tid; e < NE; e += NT) {
                const int m2 = e / K3, m3 = e % K3;
                float2 acc = plane[m2 * PS + m3];
                for (int s = 0; s < ns; ++s) {
                    const float2 B = F2[s * K2 + m2], f3 = F3[s * K3 + m3];
                    acc.x += B.x * f3.x - B.y * f3.y;
                    acc.y += B.x * f3.y + B.y * f3.x;
                }
                plane[m2 * PS + m3] = acc;
            }
            __syncthreads();
        }
    } else {
        // scatter of sum_s F1_s(m1) w2_s(x2) w3_s(x3), one lane per stencil
        // point, sites in order; then the forward DFT along axes 3 and 2
        if (tid < ord * ord) {
            const int c2 = tid / ord, c3 = tid % ord;
            for (int s = 0; s < nslt; ++s) {
                const size_t site = rowb + s;
                const int4 g = insg[site];
                const float *w = insw + site * 3 * ord;
                const float2 A = Ftab[site * per + m1];
                const int x2 = pmod(g.y - c2, K2), x3 = pmod(g.z - c3, K3);
                const float2 v = erfft::cscale(A, w[ord + c2] * w[2 * ord + c3]);
                float2 *q = plane + x2 * PS + x3;
                *q = make_float2(q->x + v.x, q->y + v.y);
            }
        }
        __syncthreads();
        lds_dft_lines<-1>(plane, K2, PS, 1, K3, t3, tid);
        __syncthreads();
        lds_dft_lines<-1>(plane, K3, 1, PS, K2, t2, tid);
        __syncthreads();
    }
    // engfac multiply + half-plane self energy (engfac is [m1][m3][m2])
    const float *ef = engfac + (size_t)m1 * NE;
    double sp = 0.0;
    for (int e = tid; e < NE; e += NT) {
        const int m2 = e % K2, m3 = e / K2;
        const float f = ef[e];
        float2 *q = plane + m2 * PS + m3;
        const float2 a = *q;
        sp += self_term(f, a);
        *q = erfft::cscale(a, f);
    }
    __syncthreads();
    lds_dft_lines<+1>(plane, K2, PS, 1, K3, t3, tid);
    __syncthreads();
    lds_dft_lines<+1>(plane, K3, 1, PS, K2, t2, tid);
    __syncthreads();
    for (int e = tid; e < NE; e += NT) Tp[e] = plane[(e / K3) * PS + (e % K3)];
    for (int off = 32; off > 0; off >>= 1) sp += __shfl_down(sp, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = sp;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int wv = 0; wv < NT / 64; ++wv) t += red[wv];
        selfpart[(size_t)n * H + m1] = plane_weight(K1, m1) * t;
    }
}

// Stage B for any K1: the complex-to-real DFT along axis 1 of one (x2, x3)
// line of NI insertions per workgroup, phi(x1) = Re X(0) + 2 sum_{m=1}^{M'}
// Re(X(m) e^{+2 pi i m x1 / K1}) [+ (-1)^x1 Re X(K1/2) for even K1], the
// unnormalised C2R transform of the Hermitian half spectrum; phi layout as
// k_pme_axis1 ([tile][x2][x3][x1][NI]).
__global__ void __launch_bounds__(ER_GEN_NT) k_pme_axis1_gen(int nchunk, int NI, int K1, int K2, int K3,
                                                             const float2 *__restrict__ T, float *__restrict__ phi,
                                                             const float2 *__restrict__ tw1)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int H = K1 / 2 + 1;
    float2 *X = reinterpret_cast<float2 *>(smem);  // [NI][H]
    float2 *tw = X + (size_t)NI * H;               // [K1]
    const int x3 = blockIdx.x, x2 = blockIdx.y, tile = blockIdx.z, tid = threadIdx.x;
    const size_t plane = (size_t)K2 * K3;
    for (int k = tid; k < NI * H; k += ER_GEN_NT) {
        const int i = k / H, m = k % H, nl = tile * NI + i;
        X[k] = (nl < nchunk) ? T[((size_t)nl * H + m) * plane + (size_t)x2 * K3 + x3] : make_float2(0.f, 0.f);
    }
    for (int k = tid; k < K1; k += ER_GEN_NT) tw[k] = tw1[k];
    __syncthreads();
    const int mh = (K1 % 2 == 0) ? H - 1 : H;  // terms m = 1 .. mh-1 appear with their conjugates
    float *dst = phi + (size_t)tile * plane * K1 * NI + ((size_t)x2 * K3 + x3) * K1 * NI;
    for (int k = tid; k < K1 * NI; k += ER_GEN_NT) {
        const int x1 = k / NI, i = k % NI;
        const float2 *Xi = X + (size_t)i * H;
        float s = 0.0f;
        int idx = 0;  // (m * x1) mod K1
        for (int m = 1; m < mh; ++m) {
            idx += x1;
            if (idx >= K1) idx -= K1;
            const float2 t = tw[idx];
            s += Xi[m].x * t.x - Xi[m].y * t.y;
        }
        s = Xi[0].x + 2.0f * s;
        if (K1 % 2 == 0) s += (x1 & 1) ? -Xi[H - 1].x : Xi[H - 1].x;
        dst[k] = s;
    }
}

// ---------------------------------------------------------------------------
// Spatial order of the molecules (per frame): counting sort of the molecule
// centres into NC^3 cells keyed by the Morton code of the cell, so
// that consecutive molecules are neighbours in space.  The gather walks the
// molecules in this order (neighbouring stencils hit the same L2 lines) and
// the pair search uses the cells as a cell list.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned int morton_spread(unsigned int v)  // 10 bits -> every third bit
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ unsigned int cell_key(int cx, int cy, int cz)
{
    return morton_spread(cx) | (morton_spread(cy) << 1) | (morton_spread(cz) << 2);
}
// fractional position of a wrapped site -> cell coordinate in [0, nc)
__device__ __forceinline__ void site_cell(const FrameConst &fc, float x, float y, float z, int nc, int &cx, int &cy,
                                          int &cz)
{
    float u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float f = fc.invP[k] * x + fc.invP[k + 3] * y + fc.invP[k + 6] * z;
        u[k] = f - floorf(f);
    }
    cx = min(nc - 1, (int)(u[0] * nc));
    cy = min(nc - 1, (int)(u[1] * nc));
    cz = min(nc - 1, (int)(u[2] * nc));
}

// the two counting passes of the frame's sorts in one launch: threads
// [0, nmol) count the molecules per cell (the molecule centre's cell), the
// next natom threads the atoms per PME brick (the brick of the spline base)
__global__ void k_sort_count(int nmol, const float4 *__restrict__ molref, FrameConst fc, int nc,
                             unsigned int *__restrict__ ccnt, unsigned int *__restrict__ keyrank, int natom,
                             const int4 *__restrict__ splg, int K1, int K2, int K3, int nbx, int nby, int bzs,
                             unsigned int *__restrict__ bcnt, unsigned int *__restrict__ akeyrank)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nmol) {
        const float4 m = molref[t];
        int cx, cy, cz;
        site_cell(fc, m.x, m.y, m.z, nc, cx, cy, cz);
        const unsigned int key = cell_key(cx, cy, cz);
        const unsigned int rank = atomicAdd(&ccnt[key], 1u);
        keyrank[2 * t] = key;
        keyrank[2 * t + 1] = rank;
        return;
    }
    const int a = t - nmol;
    if (a >= natom) return;
    const int4 g = splg[a];
    const unsigned int key = (unsigned int)((pmod(g.z, K3) / bzs * nby + pmod(g.y, K2) / ER_BRICK) * nbx +
                                            pmod(g.x, K1) / ER_BRICK);
    akeyrank[2 * a] = key;
    akeyrank[2 * a + 1] = atomicAdd(&bcnt[key], 1u);
}

// exclusive scan of n counts by one 1024-thread block -> start[n + 1]:
// each thread scans ER_SCAN_CH consecutive counts in registers, the thread
// totals go through one wave scan and one LDS exchange, so a 32768-count
// scan takes a single pass of two barriers
#define ER_SCAN_CH 32
__device__ __forceinline__ void block_scan(int ncell, const unsigned int *__restrict__ cnt,
                                           unsigned int *__restrict__ start, unsigned int *wsum)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned int carry = 0;
    for (int base = 0; base < ncell; base += 1024 * ER_SCAN_CH) {
        const int b0 = base + tid * ER_SCAN_CH;
        const bool whole = b0 + ER_SCAN_CH <= ncell;  // the thread's counts lie inside the array
        unsigned int v[ER_SCAN_CH];
        unsigned int sum = 0;
        if (whole) {  // 16-byte loads (b0 is a multiple of 32, the array 256-byte aligned)
            const uint4 *c4 = reinterpret_cast<const uint4 *>(cnt + b0);
#pragma unroll
            for (int q = 0; q < ER_SCAN_CH / 4; ++q) {
                const uint4 t = c4[q];
                v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < ER_SCAN_CH; ++k) v[k] = (b0 + k < ncell) ? cnt[b0 + k] : 0u;
        }
#pragma unroll
        for (int k = 0; k < ER_SCAN_CH; ++k) sum += v[k];
        unsigned int x = sum;  // inclusive scan of the thread totals over the wave
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        unsigned int before = 0, total = 0;
        for (int w = 0; w < 16; ++w) {
            const unsigned int t = wsum[w];
            before += (w < wave) ? t : 0u;
            total += t;
        }
        unsigned int run = carry + before + x - sum;
        if (whole) {  // start[] holds ncell + 1 entries: the 16-byte stores stay inside
            uint4 *s4 = reinterpret_cast<uint4 *>(start + b0);
#pragma unroll
            for (int q = 0; q < ER_SCAN_CH / 4; ++q) {
                uint4 t;
                t.x = run; run += v[4 * q];
                t.y = run; run += v[4 * q + 1];
                t.z = run; run += v[4 * q + 2];
                t.w = run; run += v[4 * q + 3];
                s4[q] = t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < ER_SCAN_CH; ++k) {
                if (b0 + k < ncell) start[b0 + k] = run;
                run += v[k];
            }
        }
        carry += total;
        __syncthreads();
    }
    if (tid == 0) start[ncell] = carry;
}

// both scans of the frame's sorts in one launch: block 0 the cells, block 1
// the bricks
__global__ void __launch_bounds__(1024) k_sort_scan(int ncell, const unsigned int *__restrict__ ccnt,
                                                    unsigned int *__restrict__ cstart, int nbrick,
                                                    const unsigned int *__restrict__ bcnt,
                                                    unsigned int *__restrict__ bstart)
{
    __shared__ unsigned int wsum[16];
    if (blockIdx.x == 0) block_scan(ncell, ccnt, cstart, wsum);
    else block_scan(nbrick, bcnt, bstart, wsum);
}

__global__ void k_cell_scatter(int nmol, const unsigned int *__restrict__ key_rank,
                               const unsigned int *__restrict__ start, int *__restrict__ perm)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nmol) return;
    perm[start[key_rank[2 * j]] + key_rank[2 * j + 1]] = j;
}

// ---------------------------------------------------------------------------
// PME gather (recpcal.F90:651-710 / 592-649) + residual background term
// (engproc.F90:979-1020).  A wave holds MPW = 64/LPM molecules x LPM
// insertions; with the NI-interleaved grid a load instruction reads MPW runs
// of LPM*4 bytes.  Molecules are visited in the spatial order perm[], and the
// workgroups are remapped so that each of the 8 XCDs sweeps one contiguous
// eighth of that order tile after tile (its L2 then holds one neighbourhood
// of one tile's grids).  Writes U.
// ---------------------------------------------------------------------------
template <int ORD, int NI, int LPM>
__global__ void __launch_bounds__(256) k_gather(int nmol, int nchunk, int n_off, int K1, int K2, int K3,
                                                const int *__restrict__ perm, int ngrp_x, int nxcd,
                                                const int *__restrict__ molbeg, const float4 *__restrict__ posn,
                                                const float *__restrict__ splw, const int4 *__restrict__ splg,
                                                const float *__restrict__ molq, const float *__restrict__ phi,
                                                const int *__restrict__ tag, int nslot,
                                                const SlotDesc *__restrict__ sd, float screen, float volume,
                                                float *__restrict__ U)
{
    constexpr int MPW = 64 / LPM;
    // XCD-aware remap: workgroup b runs on XCD b % nxcd; give each XCD a
    // contiguous range of molecule groups (ngrp_x is a multiple of nxcd)
    const int b = blockIdx.x;
    const int per = ngrp_x / nxcd;
    const int xcd = b % nxcd, k = b / nxcd;
    const int gx = xcd * per + (k % per);
    const int gy = k / per;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int pos = (gx * (blockDim.x >> 6) + wave) * MPW + lane / LPM;
    const int nl = gy * LPM + lane % LPM;
    if (pos >= nmol || nl >= nchunk) return;
    const int ks = ER_SLOT_OF_INS(sd, nslot, n_off + nl);
    if (n_off + nl - sd[ks].ins_off >= sd[ks].nins) return;  // padding of the solute's insertion range
    const int mol = perm[pos];
    const int tile = nl / NI, i = nl % NI;
    const size_t G = (size_t)K1 * K2 * K3;
    const float *ph = phi + (size_t)tile * G * NI + i;
    float acc = 0.0f;
    const int a0 = molbeg[mol], a1 = molbeg[mol + 1];
    for (int a = a0; a < a1; ++a) {
        const float q = posn[a].w;
        const float *w = splw + (size_t)a * 3 * ORD;
        const int4 g = splg[a];
        int i1[ORD], i2[ORD], i3[ORD];
        float w1[ORD], w2[ORD], w3[ORD];
#pragma unroll
        for (int c = 0; c < ORD; ++c) {
            i1[c] = pmod(g.x - c, K1);
            i2[c] = pmod(g.y - c, K2);
            i3[c] = pmod(g.z - c, K3);
            w1[c] = w[c];
            w2[c] = w[ORD + c];
            w3[c] = w[2 * ORD + c];
        }
        float ea = 0.0f;
#pragma unroll
        for (int c3 = 0; c3 < ORD; ++c3) {
#pragma unroll
            for (int c2 = 0; c2 < ORD; ++c2) {
                const size_t row = ((size_t)i2[c2] * K3 + i3[c3]) * K1;
                float s = 0.0f;
#pragma unroll
                for (int c1 = 0; c1 < ORD; ++c1) s += w1[c1] * ph[(row + i1[c1]) * NI];
                ea += (w3[c3] * w2[c2]) * s;
            }
        }
        // the rounded per-atom product, then the sum (as the batched path's Eat + k_mol_reduce)
        acc += __fmul_rn(q, ea);
    }
    const int n = n_off + nl;
    float qs = tag ? 0.0f : sd[ks].qslt;
    if (tag) {
        if (mol == tag[n]) acc = 0.0f;
        qs = molq[tag[n]];
    }
    if (!tag || mol != tag[n]) acc -= ER_PI_F * qs * molq[mol] / screen / screen / volume;
    const int gt = n / NI;  // n_off is a multiple of NI
    U[((size_t)gt * nmol + mol) * NI + (n % NI)] = acc;
}

// ---------------------------------------------------------------------------
// Brick gather (the batched path's form of recpcal.F90:651-710).  Atoms are
// sorted per frame by the brick (8 x 8 x BZ grid points) of their base grid
// index; a workgroup stages the (8+ORD-1)^2 x (BZ+ORD-1) region of NI
// interleaved potential grids that its brick's stencils touch in LDS once,
// then every atom of the brick reads its ORD^3 stencil from LDS (each grid
// point is read by ~15 atoms, so L2->CU traffic drops by the reuse factor).
// BZ is chosen so that two workgroups fit a CU's LDS at order 4 (one stages
// while the other computes).  Per-atom energies go to Eat and are summed per
// molecule by k_mol_reduce in the same atom order as k_gather.
// ---------------------------------------------------------------------------
// brick depth along z: order 6 takes 2 so that its staged region (13 x 13 x 7
// points x 64 bytes, 76 KB) lets two workgroups share a CU
__host__ __device__ constexpr int brick_z(int ord) { return ord <= 4 ? 6 : (ord == 6 ? 2 : 4); }

// wrap a region coordinate that lies in [-K, 2K) into [0, K)
__device__ __forceinline__ int wrap1(int x, int K)
{
    x += (x < 0) ? K : 0;
    return x - ((x >= K) ? K : 0);
}

// Per-frame atom records in brick order for the brick gather, RECF floats
// each: w1[ORD] w2[ORD] w3[ORD] q, the stencil's top corner in region
// coordinates (lx | ly << 8 | lz << 16) and the atom index (both as int
// bits), padded to whole float4s.
__host__ __device__ constexpr int brick_rec_floats(int ord) { return (3 * ord + 3 + 3) / 4 * 4; }

// Atom records in brick order; threads [natom, natom + nmol) scatter the
// molecules into cell order as k_cell_scatter does (one launch for both).
// The grid also writes bdesc[rank] = (bx | by << 10 | bz << 20, first
// atom, end atom, atom count) for the bricks in Morton order (border), so
// that the streaming gather reads one 16-byte descriptor per item and
// divides nothing (grid sides are at most 512 points, so each brick
// coordinate fits its field).
template <int ORD>
__global__ void k_brick_records(int natom, const unsigned int *__restrict__ key_rank,
                                const unsigned int *__restrict__ bstart, const float4 *__restrict__ posn,
                                const float *__restrict__ splw, const int4 *__restrict__ splg, int K1, int K2, int K3,
                                int nbx, int nby, float *__restrict__ rec, int nmol,
                                const unsigned int *__restrict__ mkeyrank, const unsigned int *__restrict__ cstart,
                                int *__restrict__ perm, int nbrick, const int *__restrict__ border,
                                uint4 *__restrict__ bdesc)
{
    constexpr int RECF = brick_rec_floats(ORD), BZ = brick_z(ORD);
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    for (int r = a; r < nbrick; r += gridDim.x * blockDim.x) {
        const int brick = border[r];
        const unsigned int bxyz = (unsigned int)(brick % nbx) | (unsigned int)((brick / nbx) % nby) << 10 |
                                  (unsigned int)(brick / (nbx * nby)) << 20;
        const unsigned int a0 = bstart[brick], a1 = bstart[brick + 1];
        bdesc[r] = make_uint4(bxyz, a0, a1, a1 - a0);
    }
    if (a >= natom) {
        const int j = a - natom;
        if (j < nmol) perm[cstart[mkeyrank[2 * j]] + mkeyrank[2 * j + 1]] = j;
        return;
    }
    const unsigned int key = key_rank[2 * a];
    const unsigned int pos = bstart[key] + key_rank[2 * a + 1];
    const int bx = key % nbx, by = (key / nbx) % nby, bz = key / (nbx * nby);
    const int4 g = splg[a];
    const int lx = pmod(g.x, K1) - bx * ER_BRICK + (ORD - 1);
    const int ly = pmod(g.y, K2) - by * ER_BRICK + (ORD - 1);
    const int lz = pmod(g.z, K3) - bz * BZ + (ORD - 1);
    float *r = rec + (size_t)pos * RECF;
    const float *w = splw + (size_t)a * 3 * ORD;
    for (int c = 0; c < 3 * ORD; ++c) r[c] = w[c];
    r[3 * ORD] = posn[a].w;
    r[3 * ORD + 1] = __int_as_float(lx | (ly << 8) | (lz << 16));
    r[3 * ORD + 2] = __int_as_float(a);
    for (int c = 3 * ORD + 3; c < RECF; ++c) r[c] = 0.f;
}

// row_ror:R, a rotation by R lanes within each 16-lane DPP row: adding
// row_ror<8> and then row_ror<4> sums the four lanes l, l+4, l+8, l+12 of a
// row into each of them whichever way the rotation turns
template <int R>
__device__ __forceinline__ float row_ror(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + R, 0xf, 0xf, false));
}

// Staging: the region is RZ * RS rows of RS grid points x NV float4, each
// row contiguous in phi (but for the periodic wrap along x); wave w stages
// rows w, w + 8, ... with one global_load_lds_dwordx4 per row (lane l ->
// point l / NV, float4 l % NV; the row's address is scalar), every load
// issued before the first wait.  Compute: sixteen lanes per atom, lane
// (c, j) owning the x offsets c1 = c, c+4, ... of the stencil and four of
// the NI = 16 insertions (one float4 of a grid point).  A 16-lane group
// reads 4 consecutive grid points x 4 float4 = 256 contiguous bytes per
// step, and a wave covers 4 atoms.  Lane (c, j) sums w1(c1) * (w3 * (w2 *
// phi)) over its column; the four x-partials are combined by two DPP row
// rotations ((c=0 + c=2) + (c=1 + c=3)).  The terms are those of k_gather;
// the association differs (FP32 rounding only).
template <int ORD, int NI>
__global__ void __launch_bounds__(512) k_gather_brick(int K1, int K2, int K3, int nbx, int nby, int nbrick,
                                                      const uint4 *__restrict__ bdesc, int nchunk,
                                                      const float *__restrict__ rec, const float *__restrict__ phi,
                                                      int natom, float *__restrict__ Eat)
{
    static_assert(NI == 16, "the brick gather stages 16 interleaved insertions");
    constexpr int BZ = brick_z(ORD);
    constexpr int RS = ER_BRICK + ORD - 1;  // region side along x and y
    constexpr int RZ = BZ + ORD - 1;        // region side along z
    constexpr int NV = NI / 4;              // float4 per grid point
    constexpr int RSU = RS * NV;            // float4 per region row
    constexpr int NROW = RS * RZ;
    constexpr int NT = 512, NW = NT / 64;
    constexpr int RPW = (NROW + NW - 1) / NW;  // rows per wave (at most)
    constexpr int XL = (ORD == 6) ? 2 : 4;  // lanes along the stencil's x offsets (6 = 2 x 3)
    constexpr int NX = (ORD + XL - 1) / XL; // x offsets per lane
    constexpr int APASS = NT / (XL * NV);   // atoms per workgroup pass
    constexpr int RECF = brick_rec_floats(ORD);
    static_assert(RSU <= 64, "a region row is staged by one wave instruction");
    extern __shared__ __align__(16) float4 reg4[];  // [RZ][RS][RS][NV]
    // XCD-aware remap: workgroup b runs on XCD b % 8.  The bricks are ranked
    // along a Morton curve (border) and each XCD owns an equal contiguous
    // range of ranks, swept tile by tile: a compact block of bricks whose
    // lower neighbours (the data a region's halo holds) were staged recently
    // on the same XCD and are still in its L2.
    const int b = blockIdx.x;
    const int per = (nbrick + 7) / 8;
    const int xcd = b & 7, k = b >> 3;
    const int rank = xcd * per + k % per, tile = k / per;
    if (rank >= nbrick || tile * NI >= nchunk) return;
    const uint4 d = bdesc[rank];  // (brick coordinates, first atom, end atom, atom count)
    const int bx = (int)(d.x & 1023u), by = (int)((d.x >> 10) & 1023u), bz = (int)(d.x >> 20);
    const unsigned int a0 = d.y, a1 = d.z;
    if (d.w == 0u) return;
    const int ox = bx * ER_BRICK - (ORD - 1), oy = by * ER_BRICK - (ORD - 1), oz = bz * BZ - (ORD - 1);
    const float4 *ph4 = reinterpret_cast<const float4 *>(phi) + (size_t)tile * K1 * K2 * K3 * NV;
    const int tid = threadIdx.x;
    const int j = tid % NV, cx = (tid / NV) % XL, s = tid / (NV * XL);
    struct AtomIn {
        int a, lx, ly, lz;
        float q, w1[NX], w2[ORD], w3[ORD];
    };
    auto load_atom = [&](unsigned int pos, AtomIn &t) {
        const float *r = rec + (size_t)pos * RECF;
#pragma unroll
        for (int m = 0; m < NX; ++m) t.w1[m] = (cx + XL * m < ORD) ? r[cx + XL * m] : 0.f;
        if constexpr (ORD == 4) {  // 64-byte records: w2, w3 and the tail are aligned float4
            const float4 v2 = *reinterpret_cast<const float4 *>(r + 4), v3 = *reinterpret_cast<const float4 *>(r + 8);
            const float4 tl = *reinterpret_cast<const float4 *>(r + 12);
            t.w2[0] = v2.x; t.w2[1] = v2.y; t.w2[2] = v2.z; t.w2[3] = v2.w;
            t.w3[0] = v3.x; t.w3[1] = v3.y; t.w3[2] = v3.z; t.w3[3] = v3.w;
            t.q = tl.x;
            const int li = __float_as_int(tl.y);
            t.lx = li & 0xff; t.ly = (li >> 8) & 0xff; t.lz = li >> 16;
            t.a = __float_as_int(tl.z);
        } else {
#pragma unroll
            for (int c = 0; c < ORD; ++c) {
                t.w2[c] = r[ORD + c];
                t.w3[c] = r[2 * ORD + c];
            }
            t.q = r[3 * ORD];
            const int li = __float_as_int(r[3 * ORD + 1]);
            t.lx = li & 0xff; t.ly = (li >> 8) & 0xff; t.lz = li >> 16;
            t.a = __float_as_int(r[3 * ORD + 2]);
        }
    };
    float4 *Eat4 = reinterpret_cast<float4 *>(Eat);
    auto compute_atom = [&](const AtomIn &t, bool live) {
        float4 ea = make_float4(0.f, 0.f, 0.f, 0.f);
        if (live) {
#pragma unroll
            for (int m = 0; m < NX; ++m) {
                // x offset c1 = cx + 4 m; its column stays inside the region
                // (lx - c1 >= 0) only for c1 < ORD, so clamp the others (w1 = 0)
                const int c1 = min(cx + XL * m, ORD - 1);
                // lowest corner of the stencil column: every read below is at
                // a constant non-negative offset from it (ds_read offset field)
                const float4 *b0 =
                    reg4 + (((t.lz - (ORD - 1)) * RS + (t.ly - (ORD - 1))) * RS + (t.lx - c1)) * NV + j;
                // order 4: all 16 reads of the column are issued before the
                // first use; larger orders read one (c3) row at a time
                constexpr int RB = (ORD <= 4) ? ORD : 1;
                float4 pr[RB][ORD];
                float4 xs = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int c3 = 0; c3 < ORD; ++c3) {
                    if (c3 % RB == 0) {
#pragma unroll
                        for (int r = 0; r < RB; ++r)
#pragma unroll
                            for (int c2 = 0; c2 < ORD; ++c2)
                                pr[r][c2] = b0[((ORD - 1 - c3 - r) * RS + (ORD - 1 - c2)) * RS * NV];
                    }
                    float4 zs = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                    for (int c2 = 0; c2 < ORD; ++c2) {
                        const float4 p = pr[c3 % RB][c2];
                        const float w2 = t.w2[c2];
                        zs.x += w2 * p.x;
                        zs.y += w2 * p.y;
                        zs.z += w2 * p.z;
                        zs.w += w2 * p.w;
                    }
                    const float w3 = t.w3[c3];
                    xs.x += w3 * zs.x;
                    xs.y += w3 * zs.y;
                    xs.z += w3 * zs.z;
                    xs.w += w3 * zs.w;
                }
                const float w1 = t.w1[m];
                ea.x += w1 * xs.x;
                ea.y += w1 * xs.y;
                ea.z += w1 * xs.z;
                ea.w += w1 * xs.w;
            }
        }
        // every lane of the wave takes part in the exchanges (the XL x-lanes
        // of an atom are XL consecutive groups of NV lanes in one DPP row)
        if constexpr (XL == 4) {
            ea.x += row_ror<8>(ea.x);
            ea.y += row_ror<8>(ea.y);
            ea.z += row_ror<8>(ea.z);
            ea.w += row_ror<8>(ea.w);
            ea.x += row_ror<4>(ea.x);
            ea.y += row_ror<4>(ea.y);
            ea.z += row_ror<4>(ea.z);
            ea.w += row_ror<4>(ea.w);
        } else {  // two x-lanes: lanes l and l ^ 4 belong to the same atom
            ea.x += __shfl_xor(ea.x, 4, 64);
            ea.y += __shfl_xor(ea.y, 4, 64);
            ea.z += __shfl_xor(ea.z, 4, 64);
            ea.w += __shfl_xor(ea.w, 4, 64);
        }
        if (live && cx == 0)
            Eat4[((size_t)tile * natom + t.a) * NV + j] = make_float4(t.q * ea.x, t.q * ea.y, t.q * ea.z, t.q * ea.w);
    };
    // the records of the first NPF passes are requested before the staging
    // loads, so that all these latencies overlap
    constexpr int NPF = 4;
    AtomIn pre[NPF];
#pragma unroll
    for (int q = 0; q < NPF; ++q)
        if (a0 + q * APASS + s < a1) load_atom(a0 + q * APASS + s, pre[q]);
    // staging straight into LDS (global_load_lds_dwordx4): lane l of a wave
    // writes its 16 bytes at the LDS address given + 16 l, so lane l of row
    // r lands at reg4[r * RSU + l]; lanes l >= RSU stay idle
    {
        const int lane = tid & 63;
        const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // scalar: row math stays in SGPRs
        const int gx = wrap1(ox + lane / NV, K1);
        const unsigned int lbyte = (unsigned int)(gx * NV + lane % NV) * 16u;
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = wv + q * NW;
            if (r < NROW) {
                const int rz = r / RS, ry = r % RS;
                const int gy = wrap1(oy + ry, K2), gz = wrap1(oz + rz, K3);
                const char *row = reinterpret_cast<const char *>(ph4 + ((size_t)gy * K3 + gz) * K1 * NV);
                if (lane < RSU)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(row + lbyte),
                        (__attribute__((address_space(3))) void *)(reg4 + r * RSU), 16, 0, 0);
            }
        }
    }
    __syncthreads();
    // pass bounds are uniform over the workgroup: all 64 lanes of a wave run
    // the DPP rotations together
#pragma unroll
    for (int q = 0; q < NPF; ++q)
        if (a0 + q * APASS < a1) compute_atom(pre[q], a0 + q * APASS + s < a1);
    for (unsigned int base = a0 + NPF * APASS; base < a1; base += APASS) {
        const unsigned int pos = base + s;
        AtomIn t;
        const bool live = pos < a1;
        if (live) load_atom(pos, t);
        compute_atom(t, live);
    }
}

// ---------------------------------------------------------------------------
// Streaming form of the brick gather (order 4): one workgroup per CU walks
// its XCD's (brick, tile) items (the same Morton-ranked ranges as
// k_gather_brick, tile after tile) with two LDS buffers, each holding one
// region and up to ER_GS_RCAP atom records.  While the atoms of item j are
// computed from one buffer, item j + W's region and records are already on
// their way into the other (global_load_lds_dwordx4), so the staging latency
// hides behind the arithmetic.  The two buffers are distinct static arrays
// and every LDS access names one of them directly, so the compiler's waits
// for the in-flight copies apply only to the buffer being filled.  Atoms
// beyond ER_GS_RCAP in one brick read their records from global memory.
// Lane mapping and arithmetic are those of k_gather_brick.
// ---------------------------------------------------------------------------
#define ER_GS_RCAP 128
#define ER_GS_NT 1024
template <int ORD, int NI>
__global__ void __launch_bounds__(ER_GS_NT) k_gather_stream(int K1, int K2, int K3, int nbx, int nby, int nbrick,
                                                             const uint4 *__restrict__ bdesc, int nchunk,
                                                             const float *__restrict__ rec,
                                                             const float *__restrict__ phi, int natom,
                                                             float *__restrict__ Eat)
{
    static_assert(NI == 16 && ORD == 4, "the streaming gather handles order 4 with 16 interleaved insertions");
    constexpr int BZ = brick_z(ORD);
    constexpr int RS = ER_BRICK + ORD - 1, RZ = BZ + ORD - 1;
    constexpr int NV = NI / 4, RSU = RS * NV, NROW = RS * RZ, NREG = NROW * RSU;
    constexpr int NT = ER_GS_NT, NW = NT / 64, RPW = (NROW + NW - 1) / NW;
    constexpr int XL = 4, APASS = NT / (XL * NV);
    constexpr int RECF = brick_rec_floats(ORD), RECU = RECF / 4;  // float4 per record
    static_assert(RECF == 16, "64-byte records");
    constexpr int NBUF = NREG + ER_GS_RCAP * RECU;
    __shared__ __align__(16) float4 bufA[NBUF];
    __shared__ __align__(16) float4 bufB[NBUF];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j4 = tid % NV, cx = (tid / NV) % XL, s = tid / (NV * XL);
    const int xcd = blockIdx.x & 7, W = gridDim.x >> 3, k0 = blockIdx.x >> 3;
    const int per = (nbrick + 7) / 8;
    const int ntile = (nchunk + NI - 1) / NI;
    const int nitem = per * ntile;
    const size_t G4 = (size_t)K1 * K2 * K3 * NV;  // float4 per tile of phi
    const float4 *rec4 = reinterpret_cast<const float4 *>(rec);
    float4 *Eat4 = reinterpret_cast<float4 *>(Eat);

    struct Item {
        int tile, bx, by, bz;
        unsigned int a0, a1;
        bool live;
    };
    // the descriptor (brick coordinates, first atom, end atom, atom count)
    // of item jj: one load, all zero (no atoms) past the XCD's bricks or the
    // items.  All four words are used, so that no register of a load in
    // flight is reused (which would wait for it) before the decode.
    auto desc = [&](int jj) {
        uint4 d = make_uint4(0u, 0u, 0u, 0u);
        if (jj < nitem) {
            const int rank = xcd * per + jj % per;
            if (rank < nbrick) d = bdesc[rank];
        }
        return d;
    };
    auto item = [&](int jj, uint4 d) {
        Item it{};
        it.tile = jj / per;
        it.bx = (int)(d.x & 1023u);
        it.by = (int)((d.x >> 10) & 1023u);
        it.bz = (int)(d.x >> 20);
        it.a0 = d.y;
        it.a1 = d.z;
        it.live = d.w > 0u;
        return it;
    };
    // issue the copies of an item's region and first records into buf
    auto stage = [&](const Item &it, float4 *buf) {
        const int ox = it.bx * ER_BRICK - (ORD - 1), oy = it.by * ER_BRICK - (ORD - 1), oz = it.bz * BZ - (ORD - 1);
        const float4 *ph4 = reinterpret_cast<const float4 *>(phi) + (size_t)it.tile * G4;
        const int gx = wrap1(ox + lane / NV, K1);
        const unsigned int lbyte = (unsigned int)(gx * NV + lane % NV) * 16u;
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = wv + q * NW;
            if (r < NROW) {
                const int rz = r / RS, ry = r % RS;
                const int gy = wrap1(oy + ry, K2), gz = wrap1(oz + rz, K3);
                const char *row = reinterpret_cast<const char *>(ph4 + ((size_t)gy * K3 + gz) * K1 * NV);
                if (lane < RSU)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(row + lbyte),
                        (__attribute__((address_space(3))) void *)(buf + r * RSU), 16, 0, 0);
            }
        }
        // records [a0, a0 + min(n, RCAP)): contiguous 64-byte rows
        const int nrec = min((int)(it.a1 - it.a0), ER_GS_RCAP);
        const int nu = nrec * RECU;
        const char *rbase = reinterpret_cast<const char *>(rec4 + (size_t)it.a0 * RECU);
        for (int u0 = wv * 64; u0 < nu; u0 += NT) {  // wave-uniform
            if (u0 + lane < nu)
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void *)(rbase + (unsigned int)(u0 + lane) * 16u),
                    (__attribute__((address_space(3))) void *)(buf + NREG + u0), 16, 0, 0);
        }
    };
    struct AtomIn {
        int a, lx, ly, lz;
        float q, w1, w2[ORD], w3[ORD];
    };
    auto unpack = [&](AtomIn &t, float w1, float4 v2, float4 v3, float4 tl) {
        t.w1 = w1;
        t.w2[0] = v2.x; t.w2[1] = v2.y; t.w2[2] = v2.z; t.w2[3] = v2.w;
        t.w3[0] = v3.x; t.w3[1] = v3.y; t.w3[2] = v3.z; t.w3[3] = v3.w;
        t.q = tl.x;
        const int li = __float_as_int(tl.y);
        t.lx = li & 0xff; t.ly = (li >> 8) & 0xff; t.lz = li >> 16;
        t.a = __float_as_int(tl.z);
    };
    // one pass of APASS atoms [base, min(base + APASS, end)) of an item from
    // buf (all lanes take part); records from LDS (FROM_LDS) or, for atoms
    // beyond the staged ER_GS_RCAP, from global memory.  The two forms stay
    // separate code so that LDS reads never wait on the copies in flight.
    auto pass = [&](const Item &it, const float4 *buf, unsigned int base, unsigned int end, auto from_lds) {
        const unsigned int pos = base + s;
        const bool live = pos < end;
        AtomIn t{};
        if (live) {
            if constexpr (decltype(from_lds)::value) {
                const float4 *r = buf + NREG + (pos - it.a0) * RECU;
                unpack(t, reinterpret_cast<const float *>(r)[cx], r[1], r[2], r[3]);
            } else {
                const float4 *r = rec4 + (size_t)pos * RECU;
                unpack(t, reinterpret_cast<const float *>(r)[cx], r[1], r[2], r[3]);
            }
        }
        float4 ea = make_float4(0.f, 0.f, 0.f, 0.f);
        if (live) {
            const float4 *b0 =
                buf + (((t.lz - (ORD - 1)) * RS + (t.ly - (ORD - 1))) * RS + (t.lx - cx)) * NV + j4;
            float4 pr[ORD][ORD];
#pragma unroll
            for (int c3 = 0; c3 < ORD; ++c3)
#pragma unroll
                for (int c2 = 0; c2 < ORD; ++c2) pr[c3][c2] = b0[((ORD - 1 - c3) * RS + (ORD - 1 - c2)) * RS * NV];
            float4 xs = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int c3 = 0; c3 < ORD; ++c3) {
                float4 zs = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int c2 = 0; c2 < ORD; ++c2) {
                    const float4 p = pr[c3][c2];
                    const float w2 = t.w2[c2];
                    zs.x += w2 * p.x;
                    zs.y += w2 * p.y;
                    zs.z += w2 * p.z;
                    zs.w += w2 * p.w;
                }
                const float w3 = t.w3[c3];
                xs.x += w3 * zs.x;
                xs.y += w3 * zs.y;
                xs.z += w3 * zs.z;
                xs.w += w3 * zs.w;
            }
            ea.x = t.w1 * xs.x;
            ea.y = t.w1 * xs.y;
            ea.z = t.w1 * xs.z;
            ea.w = t.w1 * xs.w;
        }
        ea.x += row_ror<8>(ea.x);
        ea.y += row_ror<8>(ea.y);
        ea.z += row_ror<8>(ea.z);
        ea.w += row_ror<8>(ea.w);
        ea.x += row_ror<4>(ea.x);
        ea.y += row_ror<4>(ea.y);
        ea.z += row_ror<4>(ea.z);
        ea.w += row_ror<4>(ea.w);
        if (live && cx == 0)
            Eat4[((size_t)it.tile * natom + t.a) * NV + j4] =
                make_float4(t.q * ea.x, t.q * ea.y, t.q * ea.z, t.q * ea.w);
    };
    // pass bounds are uniform over the workgroup
    auto compute = [&](const Item &it, const float4 *buf) {
        if (!it.live) return;
        const unsigned int split = min(it.a1, it.a0 + ER_GS_RCAP);
        for (unsigned int base = it.a0; base < split; base += APASS)
            pass(it, buf, base, split, std::integral_constant<bool, true>{});
        for (unsigned int base = split; base < it.a1; base += APASS)
            pass(it, buf, base, it.a1, std::integral_constant<bool, false>{});
    };

    // descriptor loads run three steps ahead (each is decoded a step after
    // it is requested, so its wait falls on a barrier instead of before the
    // staging), decoded items one step ahead of the copies, the copies one
    // step ahead of the arithmetic
    int j = k0;
    Item cur = item(j, desc(j)), nxt = item(j + W, desc(j + W));
    uint4 d2 = desc(j + 2 * W);
    if (cur.live) stage(cur, bufA);
    __syncthreads();  // (waits for the copies)
    for (;;) {
        // even step: compute from bufA, fill bufB
        uint4 d3 = desc(j + 3 * W);
        Item nx2 = item(j + 2 * W, d2);
        if (nxt.live) stage(nxt, bufB);
        compute(cur, bufA);
        __syncthreads();
        if (j + W >= nitem) break;
        j += W;
        cur = nxt;
        nxt = nx2;
        d2 = d3;
        // odd step: compute from bufB, fill bufA
        d3 = desc(j + 3 * W);
        nx2 = item(j + 2 * W, d2);
        if (nxt.live) stage(nxt, bufA);
        compute(cur, bufB);
        __syncthreads();
        if (j + W >= nitem) break;
        j += W;
        cur = nxt;
        nxt = nx2;
        d2 = d3;
    }
}

// per-molecule sum of the atom energies (atom order as in k_gather) plus the
// residual background term (engproc.F90:979-1020); writes U
#define ER_MR_MPT 8   // molecules per thread in k_mol_reduce
#define ER_MR_MPT4 2  // molecules per thread in its float4 form (16 insertions)
template <int NI>
__global__ void k_mol_reduce(int nmol, int nchunk, int n_off, int natom, const int *__restrict__ molbeg,
                             const float *__restrict__ molq, const float *__restrict__ Eat,
                             const int *__restrict__ tag, int nslot, const SlotDesc *__restrict__ sd,
                             float screen, float volume, float *__restrict__ U)
{
    if constexpr (NI == 16) {
        // block = ER_MR_MPT4 * 64 molecules x the 16 insertions of one tile
        // (blockIdx.y, one solute: solutes are padded to whole tiles);
        // thread (m, q) takes insertions 4q .. 4q + 3 of molecules m, m + 64,
        // ... as float4 (four lanes read a molecule's 64-byte atom row)
        const int q = threadIdx.x & 3, ms = threadIdx.x >> 2;
        const int tile = blockIdx.y;
        const int ks = ER_SLOT_OF_INS(sd, nslot, n_off + tile * NI);
        bool ok[4];
        int skip[4];
        float qs[4];
        bool all = true, any = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int nl = tile * NI + 4 * q + c, n = n_off + nl;
            ok[c] = nl < nchunk && n - sd[ks].ins_off < sd[ks].nins;
            skip[c] = (tag && ok[c]) ? tag[n] : -1;
            qs[c] = tag ? (ok[c] ? molq[skip[c]] : 0.0f) : sd[ks].qslt;
            all = all && ok[c];
            any = any || ok[c];
        }
        if (!any) return;
        const float4 *E4 = reinterpret_cast<const float4 *>(Eat + (size_t)tile * natom * NI) + q;
        // n_off is a whole number of tiles
        float *Ut = U + (size_t)(n_off / NI + tile) * nmol * NI + 4 * q;
        int b0[ER_MR_MPT4], b1[ER_MR_MPT4];
        float mq[ER_MR_MPT4];
        float4 e[ER_MR_MPT4][4];
#pragma unroll
        for (int r = 0; r < ER_MR_MPT4; ++r) {
            const int mol = (blockIdx.x * ER_MR_MPT4 + r) * 64 + ms;
            b0[r] = b1[r] = 0;
            mq[r] = 0.0f;
            if (mol < nmol) {
                b0[r] = molbeg[mol];
                b1[r] = molbeg[mol + 1];
                mq[r] = molq[mol];
            }
        }
#pragma unroll
        for (int r = 0; r < ER_MR_MPT4; ++r)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                e[r][k] = (b0[r] + k < b1[r]) ? E4[(size_t)(b0[r] + k) * 4] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int r = 0; r < ER_MR_MPT4; ++r) {
            const int mol = (blockIdx.x * ER_MR_MPT4 + r) * 64 + ms;
            if (mol < nmol) {
                // per insertion the same atom order and arithmetic as below
                float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (b0[r] + k < b1[r]) {
                        acc[0] += e[r][k].x;
                        acc[1] += e[r][k].y;
                        acc[2] += e[r][k].z;
                        acc[3] += e[r][k].w;
                    }
                for (int a = b0[r] + 4; a < b1[r]; ++a) {
                    const float4 v = E4[(size_t)a * 4];
                    acc[0] += v.x;
                    acc[1] += v.y;
                    acc[2] += v.z;
                    acc[3] += v.w;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (mol == skip[c]) acc[c] = 0.0f;
                    else acc[c] -= ER_PI_F * qs[c] * mq[r] / screen / screen / volume;
                }
                float *u = Ut + (size_t)mol * NI;
                if (all) {
                    *reinterpret_cast<float4 *>(u) = make_float4(acc[0], acc[1], acc[2], acc[3]);
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (ok[c]) u[c] = acc[c];
                }
            }
        }
        return;
    }
    // block = ER_MR_MPT * 256 / NI molecules x NI insertions of one tile
    // (blockIdx.y); thread (m, i) takes molecules m, m + 256 / NI, ...
    const int i = threadIdx.x % NI;
    const int mstep = blockDim.x / NI;
    const int tile = blockIdx.y;
    const int nl = tile * NI + i;
    if (nl >= nchunk) return;
    const int n = n_off + nl;
    const int ks = ER_SLOT_OF_INS(sd, nslot, n);
    if (n - sd[ks].ins_off >= sd[ks].nins) return;  // padding of the solute's insertion range
    const int skip = tag ? tag[n] : -1;
    const float qs = tag ? molq[skip] : sd[ks].qslt;
    const float *E = Eat + (size_t)tile * natom * NI + i;
    float *Urow = U + (size_t)(n / NI) * nmol * NI + (n % NI);
    // every bound, charge and the rows of molecules of up to four atoms are
    // requested before any sum, so a thread has all its loads in flight
    int b0[ER_MR_MPT], b1[ER_MR_MPT];
    float mq[ER_MR_MPT], e[ER_MR_MPT][4];
#pragma unroll
    for (int r = 0; r < ER_MR_MPT; ++r) {
        const int mol = (blockIdx.x * ER_MR_MPT + r) * mstep + threadIdx.x / NI;
        b0[r] = b1[r] = 0;
        mq[r] = 0.0f;
        if (mol < nmol) {
            b0[r] = molbeg[mol];
            b1[r] = molbeg[mol + 1];
            mq[r] = molq[mol];
        }
    }
#pragma unroll
    for (int r = 0; r < ER_MR_MPT; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) e[r][k] = (b0[r] + k < b1[r]) ? E[(size_t)(b0[r] + k) * NI] : 0.0f;
#pragma unroll
    for (int r = 0; r < ER_MR_MPT; ++r) {
        const int mol = (blockIdx.x * ER_MR_MPT + r) * mstep + threadIdx.x / NI;
        if (mol < nmol) {
            // atom order: the first four from registers, any further ones loaded here
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (b0[r] + k < b1[r]) acc += e[r][k];
            for (int a = b0[r] + 4; a < b1[r]; ++a) acc += E[(size_t)a * NI];
            if (mol == skip) acc = 0.0f;
            else acc -= ER_PI_F * qs * mq[r] / screen / screen / volume;
            Urow[(size_t)mol * NI] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// Real-space LJ + erfc-Coulomb pairs (realcal.F90:207-364 / 51-205) with
// molecule-level pruning: a molecule whose bounding-sphere centre is further
// than cutoff + R_solute + R_molecule from the solute's centre has every
// site pair beyond both cutoffs (triangle inequality on minimum images), so
// the per-pair cutoff semantics are kept exactly.  Candidates come from the
// cells within reach; each (insertion, molecule) sum is owned by one thread
// and added once to U (deterministic).
// ---------------------------------------------------------------------------
// squared distance beyond which pair_energy is exactly zero: both cutoffs
// test sqrtf(dis2) against the cutoff, so pairs with dis2 above the larger
// cutoff squared (with a relative margin for the rounding of sqrtf and of
// the square) contribute +0 and are skipped
__device__ __forceinline__ float pair_skip2(const PairParams &p)
{
    const float c = fmaxf(p.upljcut, p.elecut) * 1.0001f;
    return c * c;
}

__device__ __forceinline__ float pair_energy(const PairParams &p, float dis2, float qa, float qb, float ljeps,
                                             float ljsgm2)
{
    const float rst = sqrtf(dis2);
    float eplj = 0.0f, epcl = 0.0f;
    if (!(rst > p.upljcut)) {
        const float invr2 = ljsgm2 / dis2;
        const float invr6 = invr2 * invr2 * invr2;
        switch (p.ljswitch) {
        case 0:
        case 1:
            eplj = 4.0f * ljeps * invr6 * (invr6 - 1.0f);
            if (rst > p.lwljcut) {
                float swth;
                if (p.ljswitch == 0) {
                    const float lw2 = p.lwljcut * p.lwljcut, up2 = p.upljcut * p.upljcut;
                    const float t = dis2 - up2, den = up2 - lw2;
                    swth = (2.0f * dis2 + up2 - 3.0f * lw2) * (t * t) / (den * den * den);
                } else {
                    const float s = (rst - p.lwljcut) / (p.upljcut - p.lwljcut);
                    swth = 1.0f - 10.0f * s * s * s + 15.0f * s * s * s * s - 6.0f * s * s * s * s * s;
                }
                eplj = swth * eplj;
            }
            break;
        case 2: {
            const float sg6 = ljsgm2 * ljsgm2 * ljsgm2;
            float vdwa, vdwb;
            if (rst <= p.lwljcut) {
                vdwa = invr6 * invr6 - sg6 * sg6 / (p.lw6 * p.up6);
                vdwb = invr6 - sg6 / (p.lw3 * p.up3);
            } else {
                const float invr3 = sqrtf(invr6), sg3 = sqrtf(sg6);
                const float ta = invr6 - sg6 / p.up6, tb = invr3 - sg3 / p.up3;
                vdwa = p.up6 / (p.up6 - p.lw6) * (ta * ta);
                vdwb = p.up3 / (p.up3 - p.lw3) * (tb * tb);
            }
            eplj = 4.0f * ljeps * (vdwa - vdwb);
            break;
        }
        default: {
            const float sg6 = ljsgm2 * ljsgm2 * ljsgm2;
            float vdwa, vdwb;
            if (rst <= p.lwljcut) {
                vdwa = invr6 * invr6 - sg6 * sg6 * p.repC;
                vdwb = invr6 - sg6 * p.attC;
            } else {
                const float s = rst - p.lwljcut, s3 = s * s * s, s4 = s3 * s;
                vdwa = invr6 * invr6 - sg6 * sg6 * (p.repA * s3 + p.repB * s4 + p.repC);
                vdwb = invr6 - sg6 * (p.attA * s3 + p.attB * s4 + p.attC);
            }
            eplj = 4.0f * ljeps * (vdwa - vdwb);
        }
        }
    }
    if (!(rst >= p.elecut)) epcl = p.bare ? qa * qb / rst : qa * qb * erfcf(p.screen * rst) / rst;
    return eplj + epcl;
}

struct PairArgs {
    int nmol, nchunk, n_off, nslt, NI, ntype, nc, nsplit;
    float cell_w;  // smallest perpendicular cell width
    PairParams pp;
};

template <class LJP>
__device__ __forceinline__ float molecule_pair_sum(const PairArgs &pa, const FrameConst &fc, int n, int jm,
                                                   const int *__restrict__ molbeg, const uint4 *__restrict__ posu,
                                                   const int *__restrict__ ljtype,
                                                   const uint4 *__restrict__ insposu, LJP sltlj,
                                                   int slt_a0, bool soln, const float *__restrict__ ljene,
                                                   const float *__restrict__ ljlensq)
{
    float e = 0.0f;
    for (int s = 0; s < pa.nslt; ++s) {
        const uint4 xa = insposu[(size_t)n * pa.nslt + s];
        const int ti = soln ? ljtype[slt_a0 + s] : sltlj[s];
        for (int b = molbeg[jm]; b < molbeg[jm + 1]; ++b) {
            const uint4 xb = posu[b];
            const float d2 = lattice_d2(fc, xa, xb);
            if (d2 > pair_skip2(pa.pp)) continue;
            const int tj = ljtype[b];
            e += pair_energy(pa.pp, d2, __uint_as_float(xa.w), __uint_as_float(xb.w), ljene[ti + pa.ntype * tj],
                             ljlensq[ti + pa.ntype * tj]);
        }
    }
    return e;
}

#define ER_PAIR_CAP 2048
#define ER_PAIR_SCAP 256   // solute sites held in LDS
#define ER_PAIR_TCAP 1024  // LJ table entries (ntype^2) held in LDS

// molecule_pair_sum with the insertion's sites and the LJ tables in LDS and
// the molecule's first four atoms loaded together; same terms, same order
// Open (non-periodic) systems, bare Coulomb (realcal.F90:367-666 with
// boxshp 0, reelcut infinite): no periodic image, Coulomb between every site
// pair, LJ within upljcut.  One thread per (insertion, molecule) of the
// chunk, solute sites then molecule atoms in order; the sum is stored in U
// (every molecule is written; a soln solute's own entry is 0).
__global__ void __launch_bounds__(256) k_pairs_open(PairArgs pa, int nslot, const SlotDesc *__restrict__ sd,
                                                    const int *__restrict__ molbeg,
                                                    const float4 *__restrict__ posn, const int *__restrict__ ljtype,
                                                    const float4 *__restrict__ insposn, const int *__restrict__ tag,
                                                    const float *__restrict__ ljene,
                                                    const float *__restrict__ ljlensq, float *__restrict__ U)
{
    const int jm = blockIdx.x * blockDim.x + threadIdx.x;
    const int nl = blockIdx.y;
    if (jm >= pa.nmol || nl >= pa.nchunk) return;
    const int n = pa.n_off + nl;
    const int k = ER_SLOT_OF_INS(sd, nslot, n);
    const int local = n - sd[k].ins_off;
    PairParams p = pa.pp;
    p.elecut = __builtin_huge_valf();
    float e = 0.0f;
    const bool soln = tag != nullptr;
    if (local < sd[k].nins && !(soln && jm == tag[n])) {
        const int slt_a0 = soln ? molbeg[tag[n]] : 0;
        const int nslt = sd[k].nslt;
        const float4 *ins = insposn + sd[k].site_off + (size_t)local * nslt;
        for (int s = 0; s < nslt; ++s) {
            const float4 xa = ins[s];
            const int ti = soln ? ljtype[slt_a0 + s] : sd[k].sltlj[s];
            for (int b = molbeg[jm]; b < molbeg[jm + 1]; ++b) {
                const float4 xb = posn[b];
                const float dx = xa.x - xb.x, dy = xa.y - xb.y, dz = xa.z - xb.z;
                const int tj = ljtype[b];
                e += pair_energy(p, dx * dx + dy * dy + dz * dz, xa.w, xb.w, ljene[ti + pa.ntype * tj],
                                 ljlensq[ti + pa.ntype * tj]);
            }
        }
    }
    U[((size_t)(n / pa.NI) * pa.nmol + jm) * pa.NI + (n % pa.NI)] = e;
}

__device__ __forceinline__ float molecule_pair_sum_lds(const PairArgs &pa, const FrameConst &fc, int jm,
                                                       const int *__restrict__ molbeg,
                                                       const uint4 *__restrict__ posu,
                                                       const int *__restrict__ ljtype, const uint4 *sxu,
                                                       const int *stj, const float *te, const float *ts)
{
    const int b0 = molbeg[jm], b1 = molbeg[jm + 1];
    uint4 xb4[4];
    int tb4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        xb4[k] = (b0 + k < b1) ? posu[b0 + k] : make_uint4(0u, 0u, 0u, 0u);
        tb4[k] = (b0 + k < b1) ? ljtype[b0 + k] : 0;
    }
    float e = 0.0f;
    const float skip2 = pair_skip2(pa.pp);
    auto term = [&](const uint4 &xa, int ti, const uint4 &xb, int tj) {
        const float d2 = lattice_d2(fc, xa, xb);
        if (d2 > skip2) return;
        e += pair_energy(pa.pp, d2, __uint_as_float(xa.w), __uint_as_float(xb.w), te[ti + pa.ntype * tj],
                         ts[ti + pa.ntype * tj]);
    };
    for (int s = 0; s < pa.nslt; ++s) {
        const uint4 xa = sxu[s];
        const int ti = stj[s];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (b0 + k < b1) term(xa, ti, xb4[k], tb4[k]);
        for (int b = b0 + 4; b < b1; ++b) term(xa, ti, posu[b], ljtype[b]);
    }
    return e;
}

// The pair search and sums of one (insertion, part) block: solutes of at
// most 8 sites take one thread per candidate molecule (k_pairs), larger ones
// a group of lanes per molecule (k_pairs_many).  The two are separate kernels
// so that each gets its own register allocation: the many-site kernel is
// compiled for six waves per SIMD (the most its LDS allows), the few-site one
// needs no such limit.  A launch covers the run's blocks of both classes;
// each kernel returns at once from the other class's blocks.
template <bool MANY>
__device__ __forceinline__ void pairs_block(PairArgs pa0, int blk0, int nslot, const SlotDesc *__restrict__ sd,
    FrameConst fc, const int *__restrict__ molbeg, const uint4 *__restrict__ posu, const int *__restrict__ ljtype,
    const float4 *__restrict__ molref, const float4 *__restrict__ segref, const float *__restrict__ rmolmax,
    const unsigned int *__restrict__ cstart, const int *__restrict__ perm, const float4 *__restrict__ insposn_all,
    const uint4 *__restrict__ insposu_all, const float4 *__restrict__ insref, const int *__restrict__ tag,
    const float *__restrict__ ljene, const float *__restrict__ ljlensq, float *__restrict__ U)
{
    __shared__ int cand[ER_PAIR_CAP];
    __shared__ int ncand, cnext;
    __shared__ float4 sx[ER_PAIR_SCAP];
    __shared__ uint4 sxu[ER_PAIR_SCAP];
    __shared__ int stj[ER_PAIR_SCAP];
    __shared__ float te[ER_PAIR_TCAP], ts[ER_PAIR_TCAP];
    // workgroup b of the run (blk0 + blockIdx.x) belongs to the solute whose
    // pair_blk range holds it; pair_split workgroups per insertion, part p
    // owning the cells c with c % pair_split == p, so every (insertion,
    // molecule) sum still has a single owner
    const int b = blk0 + blockIdx.x;
    const int ks = slot_find(nslot, b, [&](int m) { return sd[m].pair_blk; });
    PairArgs pa = pa0;
    pa.nsplit = sd[ks].pair_split;
    pa.nslt = sd[ks].nslt;
    if ((pa.nslt > 8) != MANY) return;  // the other kernel's block
    const int bl = b - sd[ks].pair_blk;
    const int nl = bl / pa.nsplit, part = bl % pa.nsplit;
    const int n = sd[ks].ins_off + nl;
    const __attribute__((address_space(1))) int *sltlj = gptr(sd[ks].sltlj);
    // the insertion's sites (molecule_pair_sum reads them as insertion 0)
    const float4 *__restrict__ insposn = insposn_all + sd[ks].site_off + (size_t)nl * pa.nslt;
    const uint4 *__restrict__ insposu = insposu_all + sd[ks].site_off + (size_t)nl * pa.nslt;
    const int tid = threadIdx.x, NT = blockDim.x;
    const float4 ref = insref[n];
    const float cut = fmaxf(pa.pp.upljcut, pa.pp.elecut) + ref.w + 1e-3f;
    const bool soln = tag != nullptr;
    const int skipmol = soln ? tag[n] : -1;
    const int slt_a0 = soln ? molbeg[skipmol] : 0;
    const size_t urow = ((size_t)(n / pa.NI) * pa.nmol) * pa.NI + (n % pa.NI);
    // the insertion's sites and the LJ tables in LDS when they fit
    const int nt2 = pa.ntype * pa.ntype;
    const bool in_lds = pa.nslt <= ER_PAIR_SCAP && nt2 <= ER_PAIR_TCAP;
    if (tid == 0) ncand = 0;
    if (tid == 0) cnext = 0;  // many-site path: the next candidate a wave takes
    if (in_lds) {
        for (int k = tid; k < pa.nslt; k += NT) {
            sx[k] = insposn[k];
            sxu[k] = insposu[k];
            stj[k] = soln ? ljtype[slt_a0 + k] : sltlj[k];
        }
        for (int k = tid; k < nt2; k += NT) {
            te[k] = ljene[k];
            ts[k] = ljlensq[k];
        }
    }
    __syncthreads();
    auto pair_sum = [&](int j) {
        return in_lds ? molecule_pair_sum_lds(pa, fc, j, molbeg, posu, ljtype, sxu, stj, te, ts)
                      : molecule_pair_sum(pa, fc, 0, j, molbeg, posu, ljtype, insposu, sltlj, slt_a0, soln, ljene,
                                          ljlensq);
    };
    int cx, cy, cz;
    site_cell(fc, ref.x, ref.y, ref.z, pa.nc, cx, cy, cz);
    // a molecule centre within distance d of the search centre lies at most
    // ceil(d / width) cells away along every axis
    const int nreach = (int)ceilf((cut + rmolmax[0]) / pa.cell_w);
    const int span = min(2 * nreach + 1, pa.nc);  // distinct cells per axis
    const int lo = (span == pa.nc) ? 0 : -nreach;
    const int ncell = span * span * span;
    for (int c = part + tid * pa.nsplit; c < ncell; c += NT * pa.nsplit) {
        const int dx = c % span, dy = (c / span) % span, dz = c / (span * span);
        // cx + lo + dx lies in [-nreach, nc + nreach) with nreach < nc (or lo = 0):
        // one conditional wrap
        const int qx = wrap1(cx + lo + dx, pa.nc), qy = wrap1(cy + lo + dy, pa.nc), qz = wrap1(cz + lo + dz, pa.nc);
        const unsigned int key = cell_key(qx, qy, qz);
        const unsigned int p0 = cstart[key], p1 = cstart[key + 1];
        // the cell's molecules four at a time: indices, then centres
        for (unsigned int pb = p0; pb < p1; pb += 4) {
            int jk[4];
            float4 mk[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) jk[k] = (pb + k < p1) ? perm[pb + k] : -1;
#pragma unroll
            for (int k = 0; k < 4; ++k) mk[k] = (jk[k] >= 0) ? molref[jk[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int j = jk[k];
                if (j < 0 || j == skipmol) continue;
                const float4 m = mk[k];
                float ddx = m.x - ref.x, ddy = m.y - ref.y, ddz = m.z - ref.z;
                min_image(fc, ddx, ddy, ddz);
                const float lim = cut + m.w;
                if (ddx * ddx + ddy * ddy + ddz * ddz <= lim * lim) {
                    const int slot = atomicAdd(&ncand, 1);
                    if (slot < ER_PAIR_CAP) cand[slot] = j;
                    else  // list full: evaluate in place (same arithmetic, same single owner)
                        U[urow + (size_t)j * pa.NI] += pair_sum(j);
                }
            }
        }
    }
    __syncthreads();
    const int nc = min(ncand, ER_PAIR_CAP);
    if constexpr (!MANY) {  // few sites: one thread per candidate molecule
        for (int c = tid; c < nc; c += NT) {
            const int jm = cand[c];
            U[urow + (size_t)jm * pa.NI] += pair_sum(jm);
        }
        return;
    }
    // many sites: a group of GS lanes (16, 32 or 64) per candidate molecule;
    // lane g sums the sites g, g + GS, ... over the molecule's atoms (atom
    // order), then a fixed xor tree adds the lanes (deterministic)
    const int GS = pa.nslt <= 16 ? 16 : (pa.nslt <= 32 ? 32 : 64);
    const int cpw = 64 / GS, lane = tid & 63;
    const int grp = lane / GS, gl = lane % GS;
    const float skip2 = pair_skip2(pa.pp);
    const float skipc = fmaxf(pa.pp.upljcut, pa.pp.elecut) * 1.0001f + 1e-3f;
    auto site_sum = [&](int s, int jm) {
        const float4 xa = in_lds ? sx[s] : insposn[s];
        const uint4 xu = in_lds ? sxu[s] : insposu[s];
        const int ti = in_lds ? stj[s] : (soln ? ljtype[slt_a0 + s] : sltlj[s]);
        float e = 0.0f;
        // the molecule's centre and reach radius: beyond cutoff + reach,
        // every site pair is beyond both cutoffs (triangle inequality
        // on minimum images), so the site's sum is +0
        const float4 m = molref[jm];
        float rx = xa.x - m.x, ry = xa.y - m.y, rz = xa.z - m.z;
        min_image(fc, rx, ry, rz);
        const float lim = skipc + m.w;
        if (rx * rx + ry * ry + rz * rz > lim * lim) return e;
        // a molecule of three or more segments: the same test per segment,
        // whose skipped pairs all fail the per-pair test below (same sum,
        // same order)
        const int b0 = molbeg[jm], b1 = molbeg[jm + 1];
        const bool segmented = b1 - b0 > 2 * ER_SEG;
        for (int s0 = b0; s0 < b1; s0 += ER_SEG) {
            if (segmented) {
                const float4 q = segref[s0];
                float qx = xa.x - q.x, qy = xa.y - q.y, qz = xa.z - q.z;
                min_image(fc, qx, qy, qz);
                const float ql = skipc + q.w;
                if (qx * qx + qy * qy + qz * qz > ql * ql) continue;
            }
            for (int b = s0; b < min(s0 + ER_SEG, b1); ++b) {
                const uint4 xb = posu[b];
                const int tj = ljtype[b];
                const float d2 = lattice_d2(fc, xu, xb);
                if (d2 > skip2) continue;
                const float le = in_lds ? te[ti + pa.ntype * tj] : ljene[ti + pa.ntype * tj];
                const float ls = in_lds ? ts[ti + pa.ntype * tj] : ljlensq[ti + pa.ntype * tj];
                e += pair_energy(pa.pp, d2, xa.w, __uint_as_float(xb.w), le, ls);
            }
        }
        return e;
    };
    // a wave takes its next cpw candidates from a shared counter when it is
    // done with the last ones (molecules of very different sizes balance
    // across the waves); each candidate's sum is still one group's, in the
    // same order, so the results do not depend on which wave takes it
    int c0;
    {
        int take = 0;
        if (lane == 0) take = atomicAdd(&cnext, cpw);
        c0 = __shfl(take, 0, 64);
    }
    while (c0 < nc) {  // wave-uniform
        const int c = c0 + grp;
        const bool live = c < nc;
        const int jm = live ? cand[c] : 0;
        float e = 0.0f;
        if (live)
            for (int st = gl; st < pa.nslt; st += GS) e += site_sum(st, jm);
        for (int off = GS / 2; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
        if (live && gl == 0) U[urow + (size_t)jm * pa.NI] += e;
        int take = 0;
        if (lane == 0) take = atomicAdd(&cnext, cpw);
        c0 = __shfl(take, 0, 64);
    }
}

__global__ void __launch_bounds__(256) k_pairs(PairArgs pa0, int blk0, int nslot, const SlotDesc *__restrict__ sd,
    FrameConst fc, const int *__restrict__ molbeg, const uint4 *__restrict__ posu, const int *__restrict__ ljtype,
    const float4 *__restrict__ molref, const float4 *__restrict__ segref, const float *__restrict__ rmolmax,
    const unsigned int *__restrict__ cstart, const int *__restrict__ perm, const float4 *__restrict__ insposn_all,
    const uint4 *__restrict__ insposu_all, const float4 *__restrict__ insref, const int *__restrict__ tag,
    const float *__restrict__ ljene, const float *__restrict__ ljlensq, float *__restrict__ U)
{
    pairs_block<false>(pa0, blk0, nslot, sd, fc, molbeg, posu, ljtype, molref, segref, rmolmax, cstart, perm,
                       insposn_all, insposu_all, insref, tag, ljene, ljlensq, U);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_pairs_many(PairArgs pa0, int blk0, int nslot, const SlotDesc *__restrict__ sd,
    FrameConst fc, const int *__restrict__ molbeg, const uint4 *__restrict__ posu, const int *__restrict__ ljtype,
    const float4 *__restrict__ molref, const float4 *__restrict__ segref, const float *__restrict__ rmolmax,
    const unsigned int *__restrict__ cstart, const int *__restrict__ perm, const float4 *__restrict__ insposn_all,
    const uint4 *__restrict__ insposu_all, const float4 *__restrict__ insref, const int *__restrict__ tag,
    const float *__restrict__ ljene, const float *__restrict__ ljlensq, float *__restrict__ U)
{
    pairs_block<true>(pa0, blk0, nslot, sd, fc, molbeg, posu, ljtype, molref, segref, rmolmax, cstart, perm,
                      insposn_all, insposu_all, insref, tag, ljene, ljlensq, U);
}

// ---------------------------------------------------------------------------
// Energy-coordinate binning (update_histogram's molecule loop,
// engproc.F90:917-937, with getiduv/binsearch engproc.F90:1078-1171):
// bin = #{mesh points strictly below v} in the species' mesh.  The mesh is
// compared in FP64 in the reference; for an FP32 v, v > c holds exactly when
// v > RD32(c) (c rounded toward -inf to FP32), so the thresholds are held as
// such floats.  Grid: (NI-insertion tile) x (molecule slice); LDS
// histograms, merged into the counts with integer atomics (exact).
// Per-species min/max via order-preserving integer atomics, per-species
// energy sums per slice (reduced in a fixed order later).  The earliest
// (insertion, molecule) that hits a halt condition is kept as a key.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned int f2ord(float f)
{
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Bin brackets: the top ER_LUT_BITS of f2ord(v) select a bucket of floats;
// lut[b] = number of a species' mesh values below the bucket's smallest
// float, so every v of bucket b has its bin index (the count of mesh values
// below v) in [lut[b], lut[b + 1]], and the binary search runs over that
// bracket only (same result as over the whole species mesh).
#define ER_LUT_BITS 12
#define ER_HB 8  // energies loaded together per k_hist thread
#define ER_LUT_N ((1 << ER_LUT_BITS) + 1)

struct HistArgs {
    int nmol, nchunk, n_off, ermax, numslv, slttype, nslice;
};

// Halt key: the earliest (run, insertion, molecule) that hit a halt condition
// wins the 64-bit atomicMin.  Bits 63..42 run number, 41..21 insertion, 20..0
// molecule + 1 (0: the solute's own uvengy0).
__device__ __forceinline__ unsigned long long halt_key(unsigned int seq, int n, int jp1)
{
    return ((unsigned long long)(seq & 0x3fffffu) << 42) | ((unsigned long long)n << 21) | (unsigned long long)jp1;
}

template <int NI>
__global__ void __launch_bounds__(512) k_hist(HistArgs ha, int nslot, const SlotDesc *__restrict__ sd,
                                              const float *__restrict__ U, const int *__restrict__ tag,
                                              const float *__restrict__ meshf,
                                              const unsigned short *__restrict__ lut,
                                              const int *__restrict__ sp_begin,
                                              const int *__restrict__ sp_end, const int *__restrict__ idmin,
                                              const int *__restrict__ uvmax, const int *__restrict__ uvsoft,
                                              uint32_t *__restrict__ Cnt, double *__restrict__ flpart)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double *rsum = reinterpret_cast<double *>(smem);                 // [NT]
    float *mesh = reinterpret_cast<float *>(rsum + blockDim.x);      // [ermax]
    // [NI][hs]: an odd row stride puts one bin of the NI rows (the lanes of a
    // wave hold NI insertions) in NI different banks
    const int hs = ha.ermax | 1;
    uint32_t *hist = reinterpret_cast<uint32_t *>(mesh + ha.ermax);
    float *rmin = reinterpret_cast<float *>(hist + NI * hs);         // [NT]
    float *rmax = rmin + blockDim.x;                                  // [NT]
    unsigned short *blut = reinterpret_cast<unsigned short *>(rmax + blockDim.x);  // [numslv][ER_LUT_N]

    const int tid = threadIdx.x, NT = blockDim.x;
    const int tile = blockIdx.x, slice = blockIdx.y;
    for (int b = tid; b < ha.ermax; b += NT) mesh[b] = meshf[b];
    for (int b = tid; b < NI * hs; b += NT) hist[b] = 0u;
    for (int b = tid; b < ha.numslv * ER_LUT_N; b += NT) blut[b] = lut[b];
    __syncthreads();

    const int i = tid % NI, jj = tid / NI, jstep = NT / NI;
    const int nl = tile * NI + i;
    // the tile's solute (a tile never spans two); its padding rows are not binned
    const int ks = ER_SLOT_OF_INS(sd, nslot, ha.n_off + tile * NI);
    const int live = sd[ks].ins_off + sd[ks].nins;  // end of the solute's insertions
    const unsigned int seq = *gptr(sd[ks].seq);
    unsigned int *__restrict__ minmax = sd[ks].minmax;
    unsigned long long *__restrict__ errkey = sd[ks].errkey;
    unsigned long long *__restrict__ warnkey = sd[ks].warnkey;
    const bool valid = nl < ha.nchunk && ha.n_off + nl < live;
    const int n = ha.n_off + nl;
    const int nloc = n - sd[ks].ins_off;  // the insertion within its solute (halt keys)
    const int skip = (tag && valid) ? tag[n] : -1;
    const size_t gt = (size_t)(ha.n_off / NI + tile);
    for (int sp = 1; sp <= ha.numslv; ++sp) {
        float lmin = 3.402823466e38f, lmax = -3.402823466e38f;
        double lsum = 0.0;
        // this block's slice of the species' molecules
        const int s0 = sp_begin[sp - 1], s1 = sp_end[sp - 1];
        const int len = (s1 - s0 + ha.nslice - 1) / ha.nslice;
        const int j0 = s0 + slice * len, j1 = min(s1, j0 + len);
        const int base = idmin[sp - 1], num = uvmax[sp - 1], soft = uvsoft[sp - 1];
        if (valid) {
            // the energies of ER_HB consecutive steps are loaded together, then
            // binned in the same order as one at a time
            for (int jb = j0 + jj; jb < j1; jb += ER_HB * jstep) {
              float vb[ER_HB];
#pragma unroll
              for (int k = 0; k < ER_HB; ++k) {
                  const int j = jb + k * jstep;
                  vb[k] = (j < j1) ? U[(gt * ha.nmol + j) * NI + i] : 0.0f;
              }
#pragma unroll
              for (int k = 0; k < ER_HB; ++k) {
                const int j = jb + k * jstep;
                if (j >= j1 || j == skip) continue;
                const float v = vb[k];
                int lo = 0, hi = 0;  // NaN: bin 0 (no mesh value is below it), as the full search
                if (v == v) {
                    const unsigned short *bl = blut + (sp - 1) * ER_LUT_N + (f2ord(v) >> (32 - ER_LUT_BITS));
                    lo = bl[0];
                    hi = bl[1];
                }
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (v > mesh[base + mid]) lo = mid + 1;
                    else hi = mid;
                }
                const int idpti = lo;
                if (idpti <= 0 || (ha.slttype == 1 && idpti > soft)) gatomic_min(errkey, halt_key(seq, nloc, j + 1));
                // the 'mbin' warning (engproc.F90:1158-1167): an energy below
                // 1e3 in the open top bin; the earliest one is kept
                if (idpti == num && v < 1.0e3f) gatomic_min(warnkey, halt_key(seq, nloc, j + 1));
                atomicAdd(&hist[i * hs + base + max(idpti, 1) - 1], 1u);
                lmin = fminf(lmin, v);
                lmax = fmaxf(lmax, v);
                lsum += (double)v;
              }
            }
        }
        rmin[tid] = lmin;
        rmax[tid] = lmax;
        rsum[tid] = lsum;
        __syncthreads();
        if (tid < NI) {
            float mn = rmin[tid], mx = rmax[tid];
            double sm = rsum[tid];
            for (int k = 1; k < jstep; ++k) {
                mn = fminf(mn, rmin[tid + k * NI]);
                mx = fmaxf(mx, rmax[tid + k * NI]);
                sm += rsum[tid + k * NI];
            }
            if (tile * NI + tid < ha.nchunk && ha.n_off + tile * NI + tid < live) {
                if (mn <= mx) {
                    gatomic_min(&minmax[2 * sp], f2ord(mn));
                    gatomic_max(&minmax[2 * sp + 1], f2ord(mx));
                }
                if (flpart)
                    flpart[((size_t)(ha.n_off + tile * NI + tid) * ha.nslice + slice) * ha.numslv + sp - 1] = sm;
            }
        }
        __syncthreads();
    }
    for (int f = tid; f < NI * ha.ermax; f += NT) {
        const int ii = f / ha.ermax, b = f % ha.ermax;
        const uint32_t h = hist[ii * hs + b];
        if (h && tile * NI + ii < ha.nchunk && ha.n_off + tile * NI + ii < live)
            atomicAdd(&Cnt[(size_t)(ha.n_off + tile * NI + ii) * ha.ermax + b], h);
    }
}

__global__ void k_zero_u32(size_t n, uint32_t *__restrict__ p)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// At the end of a run: if the halt key (thread 0) or the warning key (thread
// 1) belongs to this run and has not been captured yet, record the
// offending energy while U / uvengy0 still hold it (cap[2t] = captured key,
// cap[2t + 1] = the energy's bits).
__global__ void k_halt_capture(const SlotDesc *__restrict__ sd, int NI, int nmol, const float *__restrict__ U,
                               const float *__restrict__ u0)
{
    const int t = threadIdx.x;
    if (t > 1) return;
    const SlotDesc &d = sd[blockIdx.x];  // one block per solute of the run
    const unsigned int seq = *d.seq;
    unsigned long long *cap = d.halt;
    const unsigned long long key = t == 0 ? *d.errkey : *d.warnkey;
    if (key == ~0ull || (unsigned int)(key >> 42) != (seq & 0x3fffffu) || cap[2 * t] == key) return;
    // the key holds the insertion within its solute
    const int n = d.ins_off + (int)((key >> 21) & 0x1fffffu), jp1 = (int)(key & 0x1fffffu);
    const float v = jp1 == 0 ? u0[n] : U[((size_t)(n / NI) * nmol + (jp1 - 1)) * NI + (n % NI)];
    cap[2 * t] = key;
    cap[2 * t + 1] = (unsigned long long)__float_as_uint(v);
}

// ---------------------------------------------------------------------------
// Weights (update_histogram engproc.F90:883-915 + volcorrect :1022-1038),
// uvengy0 (get_uv_energy_refs :828-855) and the scalar accumulators, for
// every insertion of a frame.
// state: [0] engnorm [1] engsmpl [2] avslf [3] voffset (double)
// flags: [0] voffset_initialized [1] usreal_set; usreal in fstate[0]
// ---------------------------------------------------------------------------
struct WeightArgs {
    int H, slttype, wgtslf, estype, selfcal, esmax, rigid, numslv, nslice;
    int bare;  // cltype 0: no reciprocal self part, no Ewald factor in volcorrect
    float w_sys, volume, screen, qtot;
    double temp;
};

// The insertions an accumulation covers.  nch < 0: every solute of the run,
// solute = blockIdx.y (k_weights) / blockIdx.x (init, fin) / the slice's
// solute (k_ecorr), each over all its insertions.  nch >= 0: solute `slot`
// only, its insertions [rel, rel + nch) (the plain path), kper insertions
// per ecorr slice.
struct AccRange {
    int slot, rel, nch, kper;
};

// (solute, first insertion in the run's space, insertion count) of a block
__device__ __forceinline__ void acc_span(const AccRange &r, const SlotDesc *__restrict__ sd, int bslot, int &k,
                                         int &n_off, int &nchunk)
{
    if (r.nch < 0) {
        k = bslot;
        n_off = sd[k].ins_off;
        nchunk = sd[k].nins;
    } else {
        k = r.slot;
        n_off = sd[k].ins_off + r.rel;
        nchunk = r.nch;
    }
}

// uvengy0 of insertion n: uvrecp (FP64 half-plane partials) + pairep +
// residual, in the reference's real arithmetic
__device__ __forceinline__ float weights_uv0(const WeightArgs &wa, const double *__restrict__ selfpart,
                                             const float *__restrict__ resself, int n, float real)
{
    double rec = 0.0;
    for (int m = 0; m < wa.H; ++m) rec += selfpart[(size_t)n * wa.H + m];
    return (float)rec + real + resself[n];
}

// first run of a block: the rigid solute's real self term and voffset (one
// block per solute)
__global__ void k_weights_init(WeightArgs wa, AccRange ar, const SlotDesc *__restrict__ sd,
                               const double *__restrict__ selfpart, const float *__restrict__ selfreal,
                               const float *__restrict__ resself)
{
    if (threadIdx.x != 0) return;
    int k, n_off, nchunk;
    acc_span(ar, sd, blockIdx.x, k, n_off, nchunk);
    int *flags = sd[k].flags;
    float *fstate = sd[k].fstate;
    if (wa.rigid && !flags[1]) {
        fstate[0] = selfreal[n_off];
        flags[1] = 1;
    }
    if (wa.wgtslf && !flags[0]) {
        const float real0 = wa.rigid ? fstate[0] : selfreal[n_off];
        sd[k].state[3] = (double)weights_uv0(wa, selfpart, resself, n_off, real0);
        flags[0] = 1;
    }
}

// per insertion (one thread each): uvengy0, the weight engnmfc
// (engproc.F90:858-912), flceng; per block: fixed-order tree partials of
// sum w, sum w_tot and the min/max of uvengy0 into wpart[block][4]
__global__ void __launch_bounds__(256) k_weights(WeightArgs wa, AccRange ar, const SlotDesc *__restrict__ sd,
                                                 const double *__restrict__ selfpart,
                                                 const float *__restrict__ selfreal,
                                                 const float *__restrict__ resself,
                                                 float *__restrict__ u0, double *__restrict__ wout,
                                                 const double *__restrict__ flpart, double *__restrict__ flceng,
                                                 double *__restrict__ wpart)
{
    __shared__ double r_w[256], r_t[256];
    __shared__ float r_mn[256], r_mx[256];
    const int tid = threadIdx.x, NT = blockDim.x;
    int k, n_off, nchunk;
    acc_span(ar, sd, blockIdx.y, k, n_off, nchunk);  // grid.y: the run's solutes
    const double voffset = gptr(sd[k].state)[3];
    const float usreal = gptr(sd[k].fstate)[0];
    const __attribute__((address_space(1))) float *wslt = sd[k].wslt ? gptr(sd[k].wslt) : nullptr;
    double sw = 0.0, st = 0.0;
    float mn = 3.402823466e38f, mx = -3.402823466e38f;
    const int nl = blockIdx.x * NT + tid;
    if (nl < nchunk) {
        const int n = n_off + nl;
        const float real = wa.rigid ? usreal : selfreal[n];
        const float u = weights_uv0(wa, selfpart, resself, n, real);
        u0[n] = u;
        mn = u;
        mx = u;
        double engnmfc = 1.0;
        if (wa.wgtslf) {
            const float factor = (float)((double)u - voffset);
            engnmfc = (wa.slttype == 1) ? exp((double)factor / wa.temp) : exp(-(double)factor / wa.temp);
        }
        float tw = wa.w_sys * (wslt ? wslt[n - sd[k].ins_off] : 1.0f);
        if (wa.estype == 2) {  // volcorrect (engproc.F90:1022-1038)
            tw = tw * wa.volume;
            if (!wa.bare) {
                const float f2 = ER_PI_F * (wa.qtot * wa.qtot) / (wa.screen * wa.screen) / wa.volume / 2.0f;
                tw = (float)((double)tw * exp((double)f2 / wa.temp));
            }
        }
        engnmfc *= (double)tw;
        wout[n] = engnmfc;
        sw = engnmfc;
        st = (double)tw;
        if (flpart) {  // flceng(sp, n): slices summed in a fixed order
            for (int sp = 0; sp < wa.numslv; ++sp) {
                double s = 0.0;
                for (int sl = 0; sl < wa.nslice; ++sl) s += flpart[((size_t)n * wa.nslice + sl) * wa.numslv + sp];
                flceng[(size_t)n * wa.numslv + sp] = s;
            }
        }
    }
    r_w[tid] = sw;
    r_t[tid] = st;
    r_mn[tid] = mn;
    r_mx[tid] = mx;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            r_w[tid] += r_w[tid + s];
            r_t[tid] += r_t[tid + s];
            r_mn[tid] = fminf(r_mn[tid], r_mn[tid + s]);
            r_mx[tid] = fmaxf(r_mx[tid], r_mx[tid + s]);
        }
        __syncthreads();
    }
    if (tid == 0) {  // each solute's partials in its own row of gridDim.x blocks
        double *wp = wpart + 4 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
        wp[0] = r_w[0];
        wp[1] = r_t[0];
        wp[2] = (double)r_mn[0];
        wp[3] = (double)r_mx[0];
    }
}

// the block partials in block order, then the sequential parts: the
// self-energy distribution and slnuv, in n as the reference adds them
__global__ void __launch_bounds__(256) k_weights_fin(WeightArgs wa, AccRange ar, const SlotDesc *__restrict__ sd,
                                                     int wrow, const double *__restrict__ wpart_all,
                                                     const float *__restrict__ u0,
                                                     const double *__restrict__ wout,
                                                     const double *__restrict__ escrd,
                                                     const double *__restrict__ flpart,
                                                     const double *__restrict__ flceng)
{
    // one block per solute; wrow = k_weights' blocks per solute row of wpart
    const int tid = threadIdx.x;
    int k, n_off, nchunk;
    acc_span(ar, sd, blockIdx.x, k, n_off, nchunk);
    const int nblk = max(1, (nchunk + 255) / 256);
    const double *__restrict__ wpart = wpart_all + 4 * (size_t)blockIdx.x * wrow;
    double *__restrict__ state = sd[k].state;
    double *__restrict__ eself = sd[k].eself;
    double *__restrict__ slnuv = sd[k].slnuv;
    unsigned int *__restrict__ minmax = sd[k].minmax;
    const unsigned int seq = *sd[k].seq;
    const int base = sd[k].ins_off;  // halt keys hold the insertion within its solute
    if (tid == 0) {
        double sw = 0.0, st = 0.0;
        float mn = 3.402823466e38f, mx = -3.402823466e38f;
        for (int b = 0; b < nblk; ++b) {
            sw += wpart[4 * b + 0];
            st += wpart[4 * b + 1];
            mn = fminf(mn, (float)wpart[4 * b + 2]);
            mx = fmaxf(mx, (float)wpart[4 * b + 3]);
        }
        state[0] += sw;
        state[1] += (double)nchunk;
        state[2] += st;
        // solute self-energy extremes (engproc.F90:914-915)
        if (nchunk > 0) {
            gatomic_min(&minmax[0], f2ord(mn));
            gatomic_max(&minmax[1], f2ord(mx));
        }
        // self-energy distribution, sequential in n as the reference adds it
        if (wa.selfcal) {
            for (int nl = 0; nl < nchunk; ++nl) {
                const int n = n_off + nl;
                const float u = u0[n];
                int lo = 0, hi = wa.esmax;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((double)u > escrd[mid]) lo = mid + 1;
                    else hi = mid;
                }
                if (lo <= 0) gatomic_min(sd[k].errkey, halt_key(seq, n - base, 0));
                if (lo == wa.esmax && u < 1.0e3f) gatomic_min(sd[k].warnkey, halt_key(seq, n - base, 0));
                gptr(eself)[max(lo, 1) - 1] += wout[n];
            }
        }
    }
    // slnuv += flceng * engnmfc (engproc.F90:939-941), sequential in n
    if (flpart && tid < wa.numslv) {
        double s = slnuv[tid];
        for (int nl = 0; nl < nchunk; ++nl) {
            const int n = n_off + nl;
            s += flceng[(size_t)n * wa.numslv + tid] * wout[n];
        }
        slnuv[tid] = s;
    }
}

// edens(b) += engnmfc * count (engproc.F90:943-947), sequential in n per bin
// edens partial of insertion slice blockIdx.y (used when no ecorr is formed
// or on the VALU ecorr path); reduced by k_ecorr_reduce with ecorr = null
// the solute and insertion range of ecorr/edens slice z (grid z or y): over
// every solute of the run (ar.nch < 0: z runs over each solute's slices in
// turn) or over one solute's range (z = the slice)
__device__ __forceinline__ void slice_span(const AccRange &ar, const SlotDesc *__restrict__ sd, int nslot, int z,
                                           int &slice, int &n_off, int &nchunk, int &kper, int &slot, int &nks)
{
    if (ar.nch < 0) {
        const int k = slot_find(nslot, z, [&](int m) { return sd[m].slice_off; });
        slice = z - sd[k].slice_off;
        n_off = sd[k].ins_off;
        nchunk = sd[k].nins;
        kper = sd[k].kper;
        slot = k;
        nks = sd[k].nks;
    } else {
        slice = z;
        n_off = sd[ar.slot].ins_off + ar.rel;
        nchunk = ar.nch;
        kper = ar.kper;
        slot = ar.slot;
        nks = (ar.nch + ar.kper - 1) / ar.kper;
    }
}

__global__ void k_edens(AccRange ar, int nslot, const SlotDesc *__restrict__ sd, int ermax,
                        const uint32_t *__restrict__ Cnt, const double *__restrict__ w, double *__restrict__ epart)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
    if (b >= ermax) return;
    int slice, n_off, nchunk, kper, slot, nks;
    slice_span(ar, sd, nslot, z, slice, n_off, nchunk, kper, slot, nks);
    const int k_begin = slice * kper, k_end = min(nchunk, k_begin + kper);
    double e = 0.0;
    for (int nl = k_begin; nl < k_end; ++nl) {
        const int n = n_off + nl;
        const uint32_t k = Cnt[(size_t)n * ermax + b];
        if (k) e += w[n] * (double)k;
    }
    epart[(size_t)z * ermax + b] = e;
}

// ---------------------------------------------------------------------------
// ecorr += C^T diag(w) C on FP64 MFMA (engproc.F90:948-962):
// ecorr(b1,b2) += sum_n w_n c_n(b1) c_n(b2).  Counts are integers and w is
// FP64, so every product is formed as the reference forms it, up to the
// summation order.  Block tile 64x64 (upper block triangle, mirrored), wave
// tile 32x32 = 2x2 MFMA 16x16x4.  The accumulator's (row, col) per
// (lane, register) comes from dmap, measured once per context by
// k_mfma_layout.
// ---------------------------------------------------------------------------
typedef double er_double4 __attribute__((ext_vector_type(4)));

// D = A B with A[i][0] = i, A[i][1] = 1, B[0][j] = 1, B[1][j] = 16 j
// (operand lane l holding A[l%16][l/16] and B[l/16][l%16]) gives
// D[i][j] = i + 16 j: each output register names its own element.
__global__ void k_mfma_layout(double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    const double a = (q == 0) ? (double)r : (q == 1 ? 1.0 : 0.0);
    const double b = (q == 0) ? 1.0 : (q == 1 ? 16.0 * r : 0.0);
    er_double4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int t = 0; t < 4; ++t) out[lane * 4 + t] = acc[t];
}

// Split-K form: workgroup (tile, slice) covers insertions [slice*kper,
// (slice+1)*kper) of the chunk and writes its 64x64 block of the partial sum
// to part[slice] (no atomics); the diagonal blocks also write the slice's
// edens partial.  k_ecorr_reduce then adds the slices in a fixed order, so
// results do not depend on scheduling.  Insertions are taken four per MFMA
// step (kq = lane / 16), the operands of the next four steps in flight
// while the current step's MFMAs run.
__global__ void __launch_bounds__(256) k_ecorr(AccRange ar, int nslot, const SlotDesc *__restrict__ sd, int ermax,
                                               int nz, const uint32_t *__restrict__ Cnt, const double *__restrict__ w,
                                               const int *__restrict__ dmap, double *__restrict__ part,
                                               double *__restrict__ epart)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // one workgroup per (slice z, upper-triangle 64 x 64 tile (by, bx)), the
    // lower triangle being mirrored.  XCD-aware order: workgroup b runs on
    // XCD b % 8, and each XCD takes one contiguous eighth of the (slice,
    // tile) list, slices outermost, so an XCD works on one solute's count
    // rows at a time and they stay in its L2.
    const int nb = (ermax + 63) / 64, ntri = nb * (nb + 1) / 2;
    const int total = ntri * nz, per = (total + 7) / 8;
    const int v = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (v >= total) return;
    const int z = v / ntri;
    int by = 0, t = v % ntri;
    while (t >= nb - by) {  // row by holds the tiles bx = by .. nb - 1
        t -= nb - by;
        ++by;
    }
    const int bx = by + t;
    int slice, n_off, nchunk, kper, slot, nks;
    slice_span(ar, sd, nslot, z, slice, n_off, nchunk, kper, slot, nks);
    // a solute with one slice adds its tile straight into ecorr and edens
    // (what k_ecorr_reduce would add: 0 + the one partial), which then
    // skips it
    const bool direct = nks == 1;
    const int b1w = by * 64 + (wave >> 1) * 32;
    const int b2w = bx * 64 + (wave & 1) * 32;
    const int k_begin = slice * kper, k_end = min(nchunk, k_begin + kper);
    const size_t plane = (size_t)ermax * ermax;
    er_double4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (er_double4){0.0, 0.0, 0.0, 0.0};
    double ed[2] = {0.0, 0.0};
    const int kq = lane >> 4, mr = lane & 15;
    auto load = [&](int k0, double (&av)[2], double (&bv)[2]) {
        const int nl = k0 + kq;
        const bool ok = nl < k_end;
        const int n = n_off + (ok ? nl : k_begin);
        const double wn = ok ? w[n] : 0.0;
        const uint32_t *row = Cnt + (size_t)n * ermax;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int b1 = b1w + a * 16 + mr;
            av[a] = (ok && b1 < ermax) ? wn * (double)row[b1] : 0.0;
            const int b2 = b2w + a * 16 + mr;
            bv[a] = (ok && b2 < ermax) ? (double)row[b2] : 0.0;
        }
    };
    // PF steps of operands in flight: stage p holds step k0 + 4 p and is
    // refilled with step k0 + 4 (p + PF) once its MFMAs are issued.  The
    // count rows come from beyond the L2 (every solute's rows are read by
    // blocks on all XCDs), so one step ahead left the MFMAs waiting; four
    // (96 registers, five waves per SIMD) measured fastest, six and eight
    // lost waves
    constexpr int PF = 4;
    double av[PF][2], bv[PF][2];
#pragma unroll
    for (int p = 0; p < PF; ++p) load(k_begin + 4 * p, av[p], bv[p]);
    for (int k0 = k_begin; k0 < k_end; k0 += 4 * PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            if (k0 + 4 * p < k_end) {
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    ed[a] += av[p][a];
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[p][a], bv[p][b], acc[a][b], 0, 0, 0);
                }
            }
            load(k0 + 4 * (p + PF), av[p], bv[p]);
        }
    }
    double *P = part + (size_t)z * plane;
    __attribute__((address_space(1))) double *E = gptr(sd[slot].ecorr);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ij = dmap[lane * 4 + r];
        const int oi = ij & 15, oj = ij >> 4;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int b1 = b1w + a * 16 + oi;
                const int b2 = b2w + b * 16 + oj;
                if (b1 < ermax && b2 < ermax) {
                    if (!direct) {
                        P[(size_t)b1 * ermax + b2] = acc[a][b][r];
                    } else if (b1 <= b2) {  // the upper triangle and its mirror, as the reduce
                        E[(size_t)b1 * ermax + b2] += acc[a][b][r];
                        if (b1 != b2) E[(size_t)b2 * ermax + b1] += acc[a][b][r];
                    }
                }
            }
    }
    // edens partial: sum the four kq groups in a fixed order (diagonal blocks,
    // left-column waves own the rows)
    if (by == bx && (wave & 1) == 0) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const double e1 = __shfl(ed[a], mr + 16, 64), e2 = __shfl(ed[a], mr + 32, 64),
                         e3 = __shfl(ed[a], mr + 48, 64);
            const double e = ((ed[a] + e1) + e2) + e3;
            const int b1 = b1w + a * 16 + mr;
            if (kq == 0 && b1 < ermax) {
                if (direct) gptr(sd[slot].edens)[b1] += e;
                else epart[(size_t)z * ermax + b1] = e;
            }
        }
    }
}

// ecorr(b1,b2) += sum over slices of part (upper triangle, mirrored);
// edens(b) += sum over slices of epart
__global__ void k_ecorr_reduce(AccRange ar, const SlotDesc *__restrict__ sd, int ermax, int full,
                               const double *__restrict__ part_all, const double *__restrict__ epart_all)
{
    // grid z: the run's solutes (ar.nch < 0) or solute ar.slot alone; each
    // adds its own slices in slice order
    const int b2 = blockIdx.x * blockDim.x + threadIdx.x, b1 = blockIdx.y;
    if (b2 >= ermax) return;
    const int ks = ar.nch < 0 ? (int)blockIdx.z : ar.slot;
    const int z0 = ar.nch < 0 ? sd[ks].slice_off : 0;
    const int nslice = ar.nch < 0 ? sd[ks].nks : (ar.nch + ar.kper - 1) / ar.kper;
    if (full && nslice == 1) return;  // k_ecorr added this solute's one slice itself
    const size_t plane = (size_t)ermax * ermax;
    const double *__restrict__ part = part_all + (size_t)z0 * plane;
    const double *__restrict__ epart = epart_all + (size_t)z0 * ermax;
    __attribute__((address_space(1))) double *ecorr = full ? gptr(sd[ks].ecorr) : nullptr;
    __attribute__((address_space(1))) double *edens = gptr(sd[ks].edens);
    if (ecorr && b1 <= b2) {
        double s = 0.0;
        for (int k = 0; k < nslice; ++k) s += part[k * plane + (size_t)b1 * ermax + b2];
        ecorr[(size_t)b1 * ermax + b2] += s;
        if (b1 != b2) ecorr[(size_t)b2 * ermax + b1] += s;
    }
    if (b1 == 0) {
        double e = 0.0;
        for (int k = 0; k < nslice; ++k) e += epart[(size_t)k * ermax + b2];
        edens[b2] += e;
    }
}

// VALU fallback of the same update (one thread per upper-triangle element)
__global__ void k_ecorr_valu(AccRange ar, const SlotDesc *__restrict__ sd, int ermax,
                             const uint32_t *__restrict__ Cnt, const double *__restrict__ w)
{
    const int b1 = blockIdx.y * 16 + threadIdx.y, b2 = blockIdx.x * 16 + threadIdx.x;
    if (b1 >= ermax || b2 >= ermax || b1 > b2) return;
    int k, n_off, nchunk;
    acc_span(ar, sd, blockIdx.z, k, n_off, nchunk);
    __attribute__((address_space(1))) double *ecorr = gptr(sd[k].ecorr);
    double s = 0.0;
    for (int nl = 0; nl < nchunk; ++nl) {
        const int n = n_off + nl;
        const uint32_t k1 = Cnt[(size_t)n * ermax + b1];
        if (!k1) continue;
        s += w[n] * (double)k1 * (double)Cnt[(size_t)n * ermax + b2];
    }
    ecorr[(size_t)b1 * ermax + b2] += s;
    if (b1 != b2) ecorr[(size_t)b2 * ermax + b1] += s;
}
