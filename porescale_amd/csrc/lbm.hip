// D3Q19 colour-gradient two-phase LBM for gfx950.
//
// Replaces the time loop body of majorityrule in
// LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f ("L1"): collision
// (L1:2407-2620), calcFs (L1:2626-2796), the bulk/interface branch and
// recolouring (L1:3701-3820), streaming (L1:3833-3839), half-way bounce-back
// (L1:3890-3915) and the density pass (L1:3934-3945); plus check_state
// (L1:2898-3296).
//
// Layout in HBM (nl fluid nodes, reference node order: x outer, z fastest):
//   populations  FP64 mode: SoA, direction-major pop[q][node] holding fr and
//                fb (2*19 doubles per node), streamed by a push through
//                dst[18][nl] (destination node of each link, -1 for a
//                bounce-back link: the population returns to the same node in
//                the opposite direction).
//                FP32 mode (lbm_fast.hip): per (q,node) T = (fr+fb) - tr_q*rr0
//                and the colour word (the minority colour's fraction of
//                fr+fb, bit 0 naming the colour), a pair of 32-bit words,
//                post-collision, in AoSoA blocks of 64 nodes (with f32x one
//                extension byte per population after them), streamed by a
//                pull through src[18][nl] (source
//                node of each incoming link, -1 for bounce-back), held on the
//                device as int16 rank differences within x-planes plus
//                per-plane bases (36 B per node instead of 72) when every
//                difference fits, as the int32 table otherwise.  T keeps the
//                small deviation of the total population (momentum is exact to
//                FP32 ulps of the deviation); the colour word (lbm_frac.h)
//                keeps the minority colour's fraction to a 25-bit significand
//                and is exact for single-colour populations, so pure-phase
//                nodes keep exactly zero minority density as the reference's
//                branch test rb<=ee requires.
//   vox[nl]      voxel index of each node in the phi grid.
//   phi grid     dense nx*ny*nz FP64, z fastest (index z + nz*(y + ny*x)):
//                colour field cf at fluid voxels, `wet` at solid voxels, so
//                the 27-point stencil of calcFs is plain loads.
//
// One time step = two kernels: collide+push (reads own populations, writes
// the streamed post-collision populations into the other buffer) and density
// (reads the new populations, writes phi).  FP64 mode compiles without FP
// contraction and keeps the reference's operation order, so its state is
// bit-identical to L1 after any number of steps.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <utility>
#include <vector>

#include "psx_common.h"

#include "lbm_device.h"
#include "lbm_fast.h"
#include "lsm_cpl.h"

using namespace psx_lbm_dev;

// ================================================================ host ====
struct psx_lbm {
    psx_lbm_params p{};
    int nx = 0, ny = 0, nz = 0;
    int64_t nl = 0, nps = 0, npb2 = 0, npb4 = 0, ns = 0;
    std::vector<int32_t> ind_raw, icor, mijk;
    std::vector<int32_t> dst_h;  // [18][nl]: push destinations (FP64) or pull sources (FP32)
    LbmConst c{};
    hipStream_t stream = nullptr;
    int32_t *d_dst = nullptr, *d_vox = nullptr;
    int16_t *d_src16 = nullptr;  // FP32 path: compressed pull table (null: use d_dst)
    int32_t *d_srca = nullptr;   // FP32 path: node-major int32 pull table [nl][20] for collide
    int32_t *d_pbase = nullptr;  // [3][nx] per-plane bases of d_src16
    uint32_t *d_etab = nullptr;  // FP32 path: the collide kernel's element table [nl][20] (null: not used)
    uint16_t *d_dtab = nullptr;  // FP32 path: the density kernel's node-major int16 pull table (null: not used)
    void *d_pop[2] = {nullptr, nullptr};
    void *d_phi = nullptr;
    double *d_partial = nullptr, *d_red = nullptr;
    int cur = 0;
    int xcd = 1;
    bool f32 = false;
    bool f32x = false;  // FP32 storage with the colour words' extension bytes
    // FP32 fused step (k_step_f32): nodes numbered brick by brick; perm maps
    // that numbering to the reference's (empty: the reference's own order)
    bool fused = false;
    int box[3] = {0, 0, 0};
    std::vector<int32_t> perm;
    psx_lbm_fast::Brick *d_bricks = nullptr;
    int nbricks = 0;
    int32_t *d_hnode = nullptr;
    int16_t *d_hcell = nullptr;
    inline int64_t ref_node(int64_t i) const { return perm.empty() ? i : perm[(size_t)i]; }
};

namespace {

inline int wrap1(int t, int n) { return t < 1 ? n : (t > n ? 1 : t); }

// Classifies link (node, q) by the rules of L1:1077-1340: returns the
// destination node id (1-based) for a streaming link, 0 for bounce-back.
// Diagonal links stream only if the destination and both single-axis
// projections are fluid (the logical intent of the product test nsum != 0).
int32_t link_dest(const psx_lbm *h, const int32_t *ic, int q) {
    auto M = [&](int x, int y, int z) {
        return h->mijk[(int64_t)(x - 1) + (int64_t)h->nx * ((y - 1) + (int64_t)h->ny * (z - 1))];
    };
    int x = wrap1(ic[0] + H_IV[q][0], h->nx), y = wrap1(ic[1] + H_IV[q][1], h->ny),
        z = wrap1(ic[2] + H_IV[q][2], h->nz);
    int32_t m = M(x, y, z);
    if (m == 0) return 0;
    if (q < 7) return m;
    int32_t m1 = M(wrap1(ic[0] + H_IV[q][0], h->nx), ic[1], ic[2]);
    int32_t m2 = M(ic[0], wrap1(ic[1] + H_IV[q][1], h->ny), ic[2]);
    int32_t m3 = M(ic[0], ic[1], wrap1(ic[2] + H_IV[q][2], h->nz));
    return (m1 != 0 && m2 != 0 && m3 != 0) ? m : 0;
}


// Node ids of the coupled model's solid lattice: corners of ind==0 voxels,
// glued in x, y and z, numbered z outer / x fastest (L1:556-645).  They fill
// columns 3.. of ipb2/ipb4 (L1:1196-1335) and give settings_19's ns.
void coupling_nodes(psx_lbm *h, std::vector<int32_t> &mijks) {
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const int ax = nx + 1, ay = ny + 1, az = nz + 1;
    std::vector<int8_t> inda((size_t)ax * ay * az, 1);
    auto IA = [&](int i, int j, int k) -> int8_t & {
        return inda[(size_t)(i - 1) + (size_t)ax * ((j - 1) + (size_t)ay * (k - 1))];
    };
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                int32_t v = h->ind_raw[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))];
                if ((v > 1 ? 1 : v) == 0)
                    for (int cc = 0; cc < 8; cc++) IA(i + (cc & 1), j + ((cc >> 1) & 1), k + (cc >> 2)) = 0;
            }
    for (int j = 1; j <= ay; j++)
        for (int k = 1; k <= az; k++) IA(1, j, k) = std::min(IA(1, j, k), IA(ax, j, k));
    for (int i = 1; i <= ax; i++)
        for (int k = 1; k <= az; k++) IA(i, 1, k) = std::min(IA(i, 1, k), IA(i, ay, k));
    for (int i = 1; i <= ax; i++)
        for (int j = 1; j <= ay; j++) IA(i, j, 1) = std::min(IA(i, j, 1), IA(i, j, az));
    mijks.assign((size_t)nx * ny * nz, 0);
    int32_t ns = 0;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                if (IA(i, j, k) == 0) mijks[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = ++ns;
    h->ns = ns;
}

void fill_const(psx_lbm *h) {
    const psx_lbm_params &p = h->p;
    LbmConst &c = h->c;
    double dt = p.dt;
    c.cr2 = p.cr2;
    c.cb2 = p.rr0 * p.cr2 / p.rb0;
    double tr1 = 1.0 - 2.0 * p.cr2 * dt * dt, tr2 = p.cr2 * dt * dt * (1.0 / 6.0),
           tr3 = p.cr2 * dt * dt * (1.0 / 12.0);
    double tb1 = 1.0 - 2.0 * c.cb2 * dt * dt, tb2 = c.cb2 * dt * dt * (1.0 / 6.0),
           tb3 = c.cb2 * dt * dt * (1.0 / 12.0);
    for (int q = 0; q < Q; q++) {
        c.tr[q] = q == 0 ? tr1 : (q < 7 ? tr2 : tr3);
        c.tb[q] = q == 0 ? tb1 : (q < 7 ? tb2 : tb3);
        c.w0[q] = c.tr[q] * p.rr0;
    }
    double svr = 1.0 / (3.0 * p.vir * dt + 0.5), svb = 1.0 / (3.0 * p.vib * dt + 0.5);
    auto trt = [&](double sv, double sq, double *S) {
        double v[Q] = {0.0, sv, sv, 0.0, sq, 0.0, sq, 0.0, sq, sv, sv, sv, sv, sv, sv, sv, sq, sq, sq};
        std::memcpy(S, v, sizeof v);
    };
    if (p.collision == PSX_COLL_MRT) {
        std::memcpy(c.Sr, p.Sr, sizeof c.Sr);
        std::memcpy(c.Sb, p.Sb, sizeof c.Sb);
    } else if (p.collision == PSX_COLL_TRT) {
        trt(svr, 1.0 / (p.lambda / (1.0 / svr - 0.5) + 0.5), c.Sr);
        trt(svb, 1.0 / (p.lambda / (1.0 / svb - 0.5) + 0.5), c.Sb);
    } else {  // collmatrTRT (L1:2858-2893)
        trt(svr, 8.0 * (2.0 - svr) / (8.0 - svr), c.Sr);
        trt(svb, 8.0 * (2.0 - svb) / (8.0 - svb), c.Sb);
    }
    double rk[Q];
    for (int q = 0; q < Q; q++) {
        rk[q] = q == 0 ? 1.0 : (q < 7 ? 1.0 / 6.0 : 1.0 / 12.0);
        double sr = 0.0, sb = 0.0;
        for (int d = 0; d < 3; d++) {
            sr += (double)H_IV[q][d] * p.bfr[d];
            sb += (double)H_IV[q][d] * p.bfb[d];
        }
        c.Fpr[q] = rk[q] * sr * (dt * dt);
        c.Fpb[q] = rk[q] * sb * (dt * dt);
    }
    for (int d = 0; d < 3; d++) { c.bfr[d] = p.bfr[d]; c.bfb[d] = p.bfb[d]; }
    c.B = p.B;
    c.sig = 0.03125 * p.sig;
    c.wet = p.wet;
    c.ee = p.ee;
    c.dt = dt;
    c.dt2 = dt * dt;
    c.dti = 1.0 / dt;
    c.nl = h->nl;
    c.nx = h->nx; c.ny = h->ny; c.nz = h->nz;
}

// extension bytes of population buffer buf (FP32 storage with extended
// colour words), else null: they follow the 2*19*aos_pad(nl) floats
uint8_t *popx(psx_lbm *h, int buf) {
    if (!h->f32x) return nullptr;
    return (uint8_t *)((float *)h->d_pop[buf] + (size_t)2 * Q * aos_pad(h->nl > 0 ? h->nl : 1));
}

int launch_density(psx_lbm *h, int buf) {
    int64_t nb = psx::ceil_div(h->nl, 256);
    if (nb == 0) return PSX_OK;
    if (h->f32) {
        PSX_HIP(psx_lbm_fast::density(h->c, (const float *)h->d_pop[buf], popx(h, buf), h->d_dst, h->d_src16,
                                      h->d_pbase, h->d_dtab, h->d_vox, (double *)h->d_phi, h->xcd, h->stream));
    } else {
        StoreF64 s{(double *)h->d_pop[buf], h->nl};
        hipLaunchKernelGGL(k_density<StoreF64>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, s,
                           h->d_vox, (double *)h->d_phi, h->xcd);
        PSX_HIP(hipGetLastError());
    }
    return PSX_OK;
}

int launch_collide(psx_lbm *h) {
    int64_t nb = psx::ceil_div(h->nl, 256);
    if (nb == 0) return PSX_OK;
    int a = h->cur, b = 1 - h->cur;
    if (h->fused) {
        PSX_HIP(psx_lbm_fast::step(h->c, (const float *)h->d_pop[a], (float *)h->d_pop[b], h->d_dst, h->d_vox,
                                   h->d_bricks, h->nbricks, h->d_hnode, h->d_hcell, h->box, h->xcd, h->stream));
    } else if (h->f32) {
        PSX_HIP(psx_lbm_fast::collide(h->c, (const float *)h->d_pop[a], popx(h, a), (float *)h->d_pop[b], popx(h, b),
                                      h->d_dst, h->d_src16, h->d_pbase, h->d_srca, h->d_etab, h->d_vox,
                                      (const double *)h->d_phi, h->xcd, h->stream));
    } else {
        StoreF64 si{(double *)h->d_pop[a], h->nl}, so{(double *)h->d_pop[b], h->nl};
        hipLaunchKernelGGL(k_collide<StoreF64>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, si, so,
                           h->d_dst, h->d_vox, (const double *)h->d_phi, h->xcd);
        PSX_HIP(hipGetLastError());
    }
    h->cur = b;
    return PSX_OK;
}

// Renumbers the FP32 nodes brick by brick (bricks of bd[0] x bd[1] x bd[2]
// voxels in x-outer / z-fastest brick order, the reference's order inside a
// brick), permutes the pull table and vox accordingly, and lists per brick
// the fluid voxels of its halo box whose colour some brick node's calcFs
// window reads.
void build_bricks(psx_lbm *h, const int (&bd)[3], std::vector<int32_t> &vox,
                  std::vector<psx_lbm_fast::Brick> &bricks, std::vector<int32_t> &hnode,
                  std::vector<int16_t> &hcell) {
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const int64_t nl = h->nl;
    const int nbx = (nx + bd[0] - 1) / bd[0], nby = (ny + bd[1] - 1) / bd[1], nbz = (nz + bd[2] - 1) / bd[2];
    const int64_t nkey = (int64_t)nbx * nby * nbz;
    auto key_of = [&](int64_t n) {
        const int32_t *ic = &h->icor[3 * n];
        return ((int64_t)((ic[0] - 1) / bd[0]) * nby + (ic[1] - 1) / bd[1]) * nbz + (ic[2] - 1) / bd[2];
    };
    std::vector<int64_t> start(nkey + 1, 0);
    for (int64_t n = 0; n < nl; n++) start[key_of(n) + 1]++;
    for (int64_t k = 0; k < nkey; k++) start[k + 1] += start[k];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    h->perm.assign((size_t)nl, 0);
    std::vector<int32_t> inv((size_t)nl);
    for (int64_t n = 0; n < nl; n++) {  // stable: reference order inside a brick
        const int64_t i = fill[key_of(n)]++;
        h->perm[(size_t)i] = (int32_t)n;
        inv[(size_t)n] = (int32_t)i;
    }
    std::vector<int32_t> src2(h->dst_h.size()), vox2((size_t)nl);
    for (int q = 1; q < Q; q++)
        for (int64_t i = 0; i < nl; i++) {
            const int32_t sidx = h->dst_h[(size_t)(q - 1) * nl + h->perm[(size_t)i]];
            src2[(size_t)(q - 1) * nl + i] = sidx >= 0 ? inv[(size_t)sidx] : -1;
        }
    for (int64_t i = 0; i < nl; i++) vox2[(size_t)i] = vox[(size_t)h->perm[(size_t)i]];
    h->dst_h.swap(src2);
    vox.swap(vox2);
    const int lx = h->box[0], ly = h->box[1], lz = h->box[2];
    std::vector<uint8_t> need((size_t)lx * ly * lz);
    std::vector<std::pair<int32_t, int16_t>> ent;
    for (int64_t k = 0; k < nkey; k++) {
        if (start[k + 1] == start[k]) continue;
        const int X = (int)(k / ((int64_t)nby * nbz)), Y = (int)((k / nbz) % nby), Z = (int)(k % nbz);
        psx_lbm_fast::Brick B{};
        B.nstart = (int32_t)start[k];
        B.ncount = (int32_t)(start[k + 1] - start[k]);
        B.x0 = X * bd[0]; B.y0 = Y * bd[1]; B.z0 = Z * bd[2];
        std::fill(need.begin(), need.end(), 0);
        for (int64_t i = start[k]; i < start[k + 1]; i++) {
            const int32_t *ic = &h->icor[3 * (size_t)h->perm[(size_t)i]];
            const int cx = ic[0] - 1 - B.x0 + 1, cy = ic[1] - 1 - B.y0 + 1, cz = ic[2] - 1 - B.z0 + 1;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++)
                    for (int e = -1; e <= 1; e++) need[(size_t)(cz + e) + lz * ((cy + b) + (size_t)ly * (cx + a))] = 1;
        }
        ent.clear();
        for (int a = 0; a < lx; a++)
            for (int b = 0; b < ly; b++)
                for (int e = 0; e < lz; e++) {
                    const size_t cell = (size_t)e + lz * (b + (size_t)ly * a);
                    if (!need[cell]) continue;
                    const int x = ((B.x0 - 1 + a) % nx + nx) % nx, y = ((B.y0 - 1 + b) % ny + ny) % ny,
                              z = ((B.z0 - 1 + e) % nz + nz) % nz;
                    const int32_t m = h->mijk[(size_t)x + (size_t)nx * (y + (size_t)ny * z)];
                    if (m > 0) ent.push_back({inv[(size_t)m - 1], (int16_t)cell});
                }
        std::sort(ent.begin(), ent.end());
        B.hstart = (int32_t)hnode.size();
        B.hcount = (int32_t)ent.size();
        for (auto &pr : ent) { hnode.push_back(pr.first); hcell.push_back(pr.second); }
        bricks.push_back(B);
    }
}

// host-side conversion of one population pair into FP32 storage: T = f - w0
// and the colour word s of the minority colour's fraction, with its
// extension byte x when ext (frac_encode_t)
inline void f32_encode(double w0, double fr, double fb, bool ext, float &T, float &s, uint8_t &x) {
    uint32_t xb = 0;
    s = ext ? frac_encode_t<8>(fr, fb, xb) : frac_encode(fr, fb);
    x = (uint8_t)xb;
    T = (float)((fr + fb) - w0);
}

}  // namespace

extern "C" int psx_lbm_create(const psx_lbm_params *p, int32_t nx, int32_t ny, int32_t nz,
                              const int32_t *ind, psx_lbm **out) {
    PSX_CHECK_ARG(p && ind && out, "psx_lbm_create: NULL argument");
    PSX_CHECK_ARG(nx >= 3 && ny >= 3 && nz >= 3, "psx_lbm_create: every dimension must be >= 3");
    PSX_CHECK_ARG((int64_t)nx * ny * nz < (int64_t)INT32_MAX, "psx_lbm_create: grid too large");
    PSX_CHECK_ARG(p->dt > 0.0 && p->rr0 > 0.0 && p->rb0 > 0.0, "psx_lbm_create: dt, rr0, rb0 must be > 0");
    PSX_CHECK_ARG(p->idir >= 1 && p->idir <= 3, "psx_lbm_create: idir must be 1, 2 or 3");
    PSX_CHECK_ARG(p->collision >= 0 && p->collision <= 2, "psx_lbm_create: unknown collision model");
    PSX_CHECK_ARG(p->precision == PSX_PREC_F64 || p->precision == PSX_PREC_F32 || p->precision == PSX_PREC_F32X,
                  "psx_lbm_create: unknown precision");
    *out = nullptr;
    int rc = psx::require_device();
    if (rc) return rc;
    PSX_GUARD_BEGIN
    psx_lbm *h = new psx_lbm();
    psx::CreateGuard<psx_lbm> guard(h, psx_lbm_destroy);
    h->p = *p;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->f32 = p->precision == PSX_PREC_F32 || p->precision == PSX_PREC_F32X;
    h->f32x = p->precision == PSX_PREC_F32X;
    if (const char *e = std::getenv("PSX_LBM_XCD")) h->xcd = std::atoi(e);  // block->XCD remap on/off
    const int64_t nv = (int64_t)nx * ny * nz;
    h->ind_raw.assign(ind, ind + nv);
    h->mijk.assign(nv, 0);
    // fluid list, x outer / z fastest, fluid iff ind == iipore (L1:1046-1072)
    int64_t nl = 0;
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) {
                int64_t a = (int64_t)(i - 1) + (int64_t)nx * ((j - 1) + (int64_t)ny * (k - 1));
                if (ind[a] == p->iipore) {
                    h->icor.push_back(i); h->icor.push_back(j); h->icor.push_back(k);
                    h->mijk[a] = (int32_t)(++nl);
                }
            }
    h->nl = nl;
    h->dst_h.assign(18 * (size_t)nl, 0);
    std::vector<int32_t> vox(nl);
    for (int64_t n = 0; n < nl; n++) {
        const int32_t *ic = &h->icor[3 * n];
        vox[n] = (int32_t)((ic[2] - 1) + (int64_t)nz * ((ic[1] - 1) + (int64_t)ny * (ic[0] - 1)));
        for (int q = 1; q < Q; q++) {
            int32_t d = link_dest(h, ic, q);
            if (d) { h->nps++; h->dst_h[(size_t)(q - 1) * nl + n] = d - 1; }
            else {
                if (q < 7) h->npb4++; else h->npb2++;
                h->dst_h[(size_t)(q - 1) * nl + n] = -1;
            }
        }
    }
    if (h->f32) {  // pull tables: source node of every incoming streaming link
        std::vector<int32_t> src(18 * (size_t)nl, -1);
        for (int q = 1; q < Q; q++)
            for (int64_t n = 0; n < nl; n++) {
                int32_t d = h->dst_h[(size_t)(q - 1) * nl + n];
                if (d >= 0) src[(size_t)(q - 1) * nl + d] = (int32_t)n;
            }
        h->dst_h.swap(src);  // the FP32 path keeps src in place of dst
    }
    std::vector<psx_lbm_fast::Brick> bricks;
    std::vector<int32_t> hnode;
    std::vector<int16_t> hcell;
    if (h->f32 && !h->f32x && nl > 0) {  // opt-in: measured slower than the two-kernel step (occupancy 2)
        const char *ef = std::getenv("PSX_LBM_FUSED");
        h->fused = ef && std::atoi(ef) != 0;
    }
    if (h->fused && (int64_t)aos_pad(nl) * 2 * Q >= (int64_t)INT32_MAX) {
        psx::set_error("psx_lbm_create: %lld fluid nodes exceed the fused step's 32-bit population index "
                       "(the two-kernel step has no such limit)", (long long)nl);
        return PSX_EINVAL;
    }
    if (h->fused) {
        int bd[3] = {8, 8, 16};  // brick (x, y, z); PSX_LBM_BRICK="bx,by,bz"
        if (const char *eb = std::getenv("PSX_LBM_BRICK")) {
            if (std::sscanf(eb, "%d,%d,%d", &bd[0], &bd[1], &bd[2]) != 3 || bd[0] < 1 || bd[1] < 1 || bd[2] < 1) {
                psx::set_error("psx_lbm_create: PSX_LBM_BRICK must be bx,by,bz");
                return PSX_EINVAL;
            }
        }
        for (int d = 0; d < 3; d++) h->box[d] = bd[d] + 2;
        if ((int64_t)h->box[0] * h->box[1] * h->box[2] * 8 > 65536) {
            psx::set_error("psx_lbm_create: brick halo box exceeds 64 KB of LDS");
            return PSX_EINVAL;
        }
        build_bricks(h, bd, vox, bricks, hnode, hcell);
    }
    // Compressed pull table (FP32 path), 36 B per node instead of 72:
    //  - direct: src = n + d with d an int16 offset; fits when every link
    //    spans fewer than 32k fluid nodes (a plane and a row for the x
    //    links, so small and thin grids);
    //  - plane-based otherwise: the source of direction q lies in plane
    //    X - c_qx; its index is n + pbase[c_qx + 1][X] + d with d the
    //    difference of the two nodes' ranks within their planes, which fits
    //    int16 on porous media (adjacent planes are alike).
    // Each form is kept only if every offset fits; PSX_LBM_IDX16=0 keeps the
    // int32 table only, PSX_LBM_IDX16=2 skips the direct form.
    std::vector<int16_t> s16;
    std::vector<int32_t> pbase;
    bool use16 = false, direct = false;
    int idx16 = 1;
    if (h->f32 && !h->fused && nl > 0) {
        const char *e16 = std::getenv("PSX_LBM_IDX16");
        idx16 = e16 ? std::atoi(e16) : 1;
        use16 = idx16 != 0;
    }
    if (use16 && idx16 == 1) {
        direct = true;
        s16.assign(18 * (size_t)nl, 0);
        for (int q = 1; q < Q && direct; q++)
            for (int64_t n = 0; n < nl; n++) {
                const int32_t sidx = h->dst_h[(size_t)(q - 1) * nl + n];
                if (sidx < 0) { s16[(size_t)(q - 1) * nl + n] = (int16_t)-32768; continue; }
                const int64_t d = (int64_t)sidx - n;
                if (d < -32767 || d > 32767) { direct = false; break; }
                s16[(size_t)(q - 1) * nl + n] = (int16_t)d;
            }
    }
    if (use16 && !direct) {
        std::vector<int64_t> pstart(nx + 1, 0);
        for (int64_t n = 0; n < nl; n++) pstart[h->icor[3 * n]]++;  // counts of plane x-1 at index x
        for (int X = 0; X < nx; X++) pstart[X + 1] += pstart[X];
        pbase.assign(3 * (size_t)nx, 0);
        for (int cx = -1; cx <= 1; cx++)
            for (int X = 0; X < nx; X++)
                pbase[(size_t)(cx + 1) * nx + X] = (int32_t)(pstart[((X - cx) % nx + nx) % nx] - pstart[X]);
        s16.assign(18 * (size_t)nl, 0);
        for (int q = 1; q < Q && use16; q++) {
            const int cx = H_IV[q][0];
            for (int64_t n = 0; n < nl; n++) {
                int32_t sidx = h->dst_h[(size_t)(q - 1) * nl + n];
                if (sidx < 0) { s16[(size_t)(q - 1) * nl + n] = (int16_t)-32768; continue; }
                const int X = h->icor[3 * n] - 1;
                int64_t d = (int64_t)sidx - n - pbase[(size_t)(cx + 1) * nx + X];
                if (d < -32767 || d > 32767) { use16 = false; break; }
                s16[(size_t)(q - 1) * nl + n] = (int16_t)d;
            }
        }
    }
    fill_const(h);
    PSX_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    size_t popb = h->f32 ? (size_t)aos_pad(nl > 0 ? nl : 1) * 2 * Q * sizeof(float) +
                               (h->f32x ? (size_t)xbytes(nl > 0 ? nl : 1) : 0)
                         : (size_t)(nl > 0 ? nl : 1) * 2 * Q * sizeof(double);
    PSX_HIP(hipMalloc(&h->d_pop[0], popb));
    PSX_HIP(hipMalloc(&h->d_pop[1], popb));
    PSX_HIP(hipMemsetAsync(h->d_pop[0], 0, popb, h->stream));
    PSX_HIP(hipMalloc(&h->d_dst, sizeof(int32_t) * 18 * (size_t)(nl > 0 ? nl : 1)));
    PSX_HIP(hipMalloc(&h->d_vox, sizeof(int32_t) * (size_t)(nl > 0 ? nl : 1)));
    // With the extended colour words a collide wavefront would hold more
    // loads than it can keep in flight (63) with 18 table loads per node:
    // its int32 table, taken when the direct int16 form does not fit, goes
    // node-major, five 16-byte loads a node (measured at 256^3 with T and the
    // word in separate slots, three loads per population: collide 0.61 ms
    // instead of 0.92; without the extension it gains nothing, 0.567 vs
    // 0.562 ms; PSX_LBM_SRCAOS=0/1 forces it off/on):
    const char *ea = std::getenv("PSX_LBM_SRCAOS");
    const bool node_major = ea ? std::atoi(ea) != 0 : h->f32x;
    // The collide kernel's element table (collide_pull_e): per node the 18
    // pulled AoSoA element indices with bounce-back resolved, the voxel and
    // the packed coordinates, so the kernel forms no address and divides
    // nothing (collide 2-4 % faster at 256^3 than with the int32 source
    // table).  The density kernel keeps its int16 table: reading this one
    // instead (80 B a node for 40) made it 29 % slower.  It needs element
    // byte offsets below 2^32 (a population buffer below 4 GiB, 28 M fluid
    // nodes) and dimensions below 1025; PSX_LBM_ETAB=0 keeps the index
    // tables above.
    const char *ee = std::getenv("PSX_LBM_ETAB");
    const bool etab = h->f32 && !h->fused && nl > 0 && (!ee || std::atoi(ee) != 0) && !psx_lbm_fast::wide_index(nl) &&
                      (int64_t)Q * aos_pad(nl) * 8 <= (int64_t)UINT32_MAX && nx <= 1024 && ny <= 1024 && nz <= 1024;
    if (etab) {
        std::vector<uint32_t> et(20 * (size_t)nl);
        for (int64_t n = 0; n < nl; n++) {
            uint32_t *row = &et[20 * (size_t)n];
            for (int q = 1; q < Q; q++) {
                const int32_t sidx = h->dst_h[(size_t)(q - 1) * nl + n];
                row[q - 1] = (uint32_t)(sidx >= 0 ? aos(q, sidx) : aos(opp(q), n));
            }
            const int32_t *ic = &h->icor[3 * n];
            row[18] = (uint32_t)vox[n];
            row[19] = (uint32_t)(ic[0] - 1) | (uint32_t)(ic[1] - 1) << 10 | (uint32_t)(ic[2] - 1) << 20;
        }
        PSX_HIP(hipMalloc(&h->d_etab, sizeof(uint32_t) * et.size()));
        PSX_HIP(hipMemcpyAsync(h->d_etab, et.data(), sizeof(uint32_t) * et.size(), hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    }
    if (h->f32 && !h->fused && nl > 0 && !(use16 && direct) && !etab) {
        if (node_major) {
            std::vector<int32_t> sa(20 * (size_t)nl, -1);
            for (int q = 1; q < Q; q++)
                for (int64_t n = 0; n < nl; n++) sa[20 * (size_t)n + (q - 1)] = h->dst_h[(size_t)(q - 1) * nl + n];
            PSX_HIP(hipMalloc(&h->d_srca, sizeof(int32_t) * sa.size()));
            PSX_HIP(hipMemcpyAsync(h->d_srca, sa.data(), sizeof(int32_t) * sa.size(), hipMemcpyHostToDevice, h->stream));
            PSX_HIP(hipStreamSynchronize(h->stream));
        }
    }
    // With the element table the int16 pull table serves the density kernel
    // alone; it goes node-major there, the 18 offsets of a node in 32 + 4 B
    // and its voxel beside the last two: three loads a node instead of 19.
    const char *ed = std::getenv("PSX_LBM_DTAB");
    if (use16 && etab && (!ed || std::atoi(ed) != 0)) {
        std::vector<uint16_t> dt(20 * (size_t)nl);
        for (int64_t n = 0; n < nl; n++) {
            for (int q = 1; q <= 16; q++) dt[16 * (size_t)n + q - 1] = (uint16_t)s16[(size_t)(q - 1) * nl + n];
            uint16_t *t = &dt[16 * (size_t)nl + 4 * (size_t)n];
            t[0] = (uint16_t)s16[(size_t)16 * nl + n];
            t[1] = (uint16_t)s16[(size_t)17 * nl + n];
            const uint32_t v = (uint32_t)vox[n];
            std::memcpy(&t[2], &v, 4);
        }
        PSX_HIP(hipMalloc(&h->d_dtab, sizeof(uint16_t) * dt.size()));
        PSX_HIP(hipMemcpyAsync(h->d_dtab, dt.data(), sizeof(uint16_t) * dt.size(), hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    }
    if (use16) {
        if (!h->d_dtab) {  // else no kernel reads the direction-major form
            PSX_HIP(hipMalloc(&h->d_src16, sizeof(int16_t) * s16.size()));
            PSX_HIP(hipMemcpyAsync(h->d_src16, s16.data(), sizeof(int16_t) * s16.size(), hipMemcpyHostToDevice,
                                   h->stream));
        }
        if (!direct) {  // the plane-based form needs the bases
            PSX_HIP(hipMalloc(&h->d_pbase, sizeof(int32_t) * 3 * (size_t)nx));
            PSX_HIP(hipMemcpyAsync(h->d_pbase, pbase.data(), sizeof(int32_t) * 3 * nx, hipMemcpyHostToDevice, h->stream));
        }
    }
    size_t phib = (size_t)nv * sizeof(double);
    PSX_HIP(hipMalloc(&h->d_phi, phib));
    int64_t nb = std::max<int64_t>(1, psx::ceil_div(nl, 256));
    PSX_HIP(hipMalloc(&h->d_partial, sizeof(double) * NDIAG * nb));
    PSX_HIP(hipMalloc(&h->d_red, sizeof(double) * NDIAG));
    if (h->fused) {
        h->nbricks = (int)bricks.size();
        PSX_HIP(hipMalloc(&h->d_bricks, sizeof(psx_lbm_fast::Brick) * std::max<size_t>(1, bricks.size())));
        PSX_HIP(hipMalloc(&h->d_hnode, sizeof(int32_t) * std::max<size_t>(1, hnode.size())));
        PSX_HIP(hipMalloc(&h->d_hcell, sizeof(int16_t) * std::max<size_t>(1, hcell.size())));
        PSX_HIP(hipMemcpyAsync(h->d_bricks, bricks.data(), sizeof(psx_lbm_fast::Brick) * bricks.size(),
                               hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipMemcpyAsync(h->d_hnode, hnode.data(), sizeof(int32_t) * hnode.size(), hipMemcpyHostToDevice,
                               h->stream));
        PSX_HIP(hipMemcpyAsync(h->d_hcell, hcell.data(), sizeof(int16_t) * hcell.size(), hipMemcpyHostToDevice,
                               h->stream));
    }
    if (nl > 0) {
        PSX_HIP(hipMemcpyAsync(h->d_dst, h->dst_h.data(), sizeof(int32_t) * 18 * nl, hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipMemcpyAsync(h->d_vox, vox.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, h->stream));
    }
    int64_t gb = psx::ceil_div(nv, 256);
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)gb), dim3(256), 0, h->stream, (double *)h->d_phi, nv, p->wet);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipStreamSynchronize(h->stream));
    *out = guard.release();
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" void psx_lbm_destroy(psx_lbm *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *bufs[] = {h->d_pop[0], h->d_pop[1], h->d_dst, h->d_vox, h->d_phi, h->d_partial, h->d_red, h->d_src16,
                    h->d_pbase, h->d_bricks, h->d_hnode, h->d_hcell, h->d_srca, h->d_etab, h->d_dtab};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int psx_lbm_counts(psx_lbm *h, int64_t *counts) {
    PSX_CHECK_ARG(h && counts, "psx_lbm_counts: NULL argument");
    // the coupled model's solid node count (settings_19) comes with the tables
    if (h->ns == 0) {
        PSX_GUARD_BEGIN
        std::vector<int32_t> mijks;
        coupling_nodes(h, mijks);
        PSX_GUARD_END
    }
    counts[0] = h->nl; counts[1] = h->nps; counts[2] = h->npb2; counts[3] = h->npb4;
    counts[4] = h->ns;
    return PSX_OK;
}

extern "C" int psx_lbm_tables(psx_lbm *h, int32_t *icor, int32_t *ips, int32_t *ipb2, int32_t *ipb4) {
    PSX_CHECK_ARG(h, "psx_lbm_tables: NULL handle");
    PSX_GUARD_BEGIN
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    if (icor) std::memcpy(icor, h->icor.data(), sizeof(int32_t) * 3 * h->nl);
    if (!ips && !ipb2 && !ipb4) return PSX_OK;
    std::vector<int32_t> mijks;
    coupling_nodes(h, mijks);
    auto MS = [&](int x, int y, int z) {
        return mijks[(size_t)(wrap1(x, nx) - 1) + (size_t)nx * ((wrap1(y, ny) - 1) + (size_t)ny * (wrap1(z, nz) - 1))];
    };
    int64_t s = 0, b2 = 0, b4 = 0;
    for (int64_t n = 0; n < h->nl; n++) {
        const int32_t *ic = &h->icor[3 * n];
        for (int q = 1; q < Q; q++) {
            int32_t d = link_dest(h, ic, q) - 1;
            const int *v = H_IV[q];
            if (d >= 0) {
                if (ips) { ips[3 * s] = (int32_t)(n + 1); ips[3 * s + 1] = d + 1; ips[3 * s + 2] = q + 1; }
                s++;
            } else if (q < 7) {
                if (ipb4) {
                    int32_t *e = &ipb4[6 * b4];
                    e[0] = (int32_t)(n + 1); e[1] = opp(q) + 1;
                    for (int m = 0; m < 4; m++) {
                        int o1 = m & 1, o2 = (m >> 1) & 1, off[3];
                        if (v[0] != 0) { off[0] = (v[0] + 1) / 2; off[1] = o1; off[2] = o2; }
                        else if (v[1] != 0) { off[0] = o1; off[1] = (v[1] + 1) / 2; off[2] = o2; }
                        else { off[0] = o1; off[1] = o2; off[2] = (v[2] + 1) / 2; }
                        e[2 + m] = MS(ic[0] + off[0], ic[1] + off[1], ic[2] + off[2]);
                    }
                }
                b4++;
            } else {
                if (ipb2) {
                    int32_t *e = &ipb2[4 * b2];
                    e[0] = (int32_t)(n + 1); e[1] = opp(q) + 1;
                    for (int m = 0; m < 2; m++) {
                        int off[3];
                        if (v[0] == 0) { off[0] = m; off[1] = (v[1] + 1) / 2; off[2] = (v[2] + 1) / 2; }
                        else if (v[1] == 0) { off[0] = (v[0] + 1) / 2; off[1] = m; off[2] = (v[2] + 1) / 2; }
                        else { off[0] = (v[0] + 1) / 2; off[1] = (v[1] + 1) / 2; off[2] = m; }
                        e[2 + m] = MS(ic[0] + off[0], ic[1] + off[1], ic[2] + off[2]);
                    }
                }
                b2++;
            }
        }
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_rates(psx_lbm *h, double *Sr, double *Sb) {
    PSX_CHECK_ARG(h && Sr && Sb, "psx_lbm_rates: NULL argument");
    std::memcpy(Sr, h->c.Sr, sizeof h->c.Sr);
    std::memcpy(Sb, h->c.Sb, sizeof h->c.Sb);
    return PSX_OK;
}

extern "C" int psx_lbm_set_state(psx_lbm *h, const double *fr, const double *fb) {
    PSX_CHECK_ARG(h && fr && fb, "psx_lbm_set_state: NULL argument");
    if (h->nl == 0) return PSX_OK;
    PSX_GUARD_BEGIN
    const int64_t nl = h->nl;
    h->cur = 0;
    if (h->f32) {
        // stored state is post-collision and not yet streamed: place f_q(i)
        // where the pull of node i reads it (inverse of the streaming step)
        std::vector<float> st((size_t)2 * Q * aos_pad(nl), 0.0f);
        std::vector<uint8_t> sx(h->f32x ? (size_t)xbytes(nl) : 1, 0);
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++) {
                int64_t j = i;
                int k = q;
                if (q > 0) {
                    int32_t sidx = h->dst_h[(size_t)(q - 1) * nl + i];
                    if (sidx >= 0) j = sidx;
                    else k = opp(q);
                }
                const int64_t r = h->ref_node(i);
                f32_encode(h->c.w0[q], fr[Q * r + q], fb[Q * r + q], h->f32x, st[2 * aos(k, j)], st[2 * aos(k, j) + 1],
                           sx[h->f32x ? aosx(k, j) : 0]);
            }
        PSX_HIP(hipMemcpyAsync(h->d_pop[0], st.data(), sizeof(float) * st.size(), hipMemcpyHostToDevice, h->stream));
        if (h->f32x)
            PSX_HIP(hipMemcpyAsync(popx(h, 0), sx.data(), sx.size(), hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    } else {
        std::vector<double> st((size_t)2 * Q * nl);
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++) {
                st[(size_t)q * nl + i] = fr[Q * i + q];
                st[(size_t)(Q + q) * nl + i] = fb[Q * i + q];
            }
        PSX_HIP(hipMemcpyAsync(h->d_pop[0], st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    }
    int rc = launch_density(h, 0);
    if (rc) return rc;
    PSX_HIP(hipStreamSynchronize(h->stream));
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_get_state(psx_lbm *h, double *fr, double *fb) {
    PSX_CHECK_ARG(h && fr && fb, "psx_lbm_get_state: NULL argument");
    if (h->nl == 0) return PSX_OK;
    PSX_GUARD_BEGIN
    const int64_t nl = h->nl;
    if (h->f32) {
        std::vector<float> st((size_t)2 * Q * aos_pad(nl));
        std::vector<uint8_t> sx(h->f32x ? (size_t)xbytes(nl) : 0);
        PSX_HIP(hipMemcpyAsync(st.data(), h->d_pop[h->cur], sizeof(float) * st.size(), hipMemcpyDeviceToHost, h->stream));
        if (h->f32x)
            PSX_HIP(hipMemcpyAsync(sx.data(), popx(h, h->cur), sx.size(), hipMemcpyDeviceToHost, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++) {
                int64_t j = i;
                int k = q;
                if (q > 0) {
                    int32_t sidx = h->dst_h[(size_t)(q - 1) * nl + i];
                    if (sidx >= 0) j = sidx;
                    else k = opp(q);
                }
                double f = h->c.w0[q] + (double)st[2 * aos(k, j)];
                const int64_t o = h->ref_node(i);
                frac_split(f, st[2 * aos(k, j) + 1], fr[Q * o + q], fb[Q * o + q], h->f32x ? sx[aosx(k, j)] : 0u);
            }
    } else {
        std::vector<double> st((size_t)2 * Q * nl);
        PSX_HIP(hipMemcpyAsync(st.data(), h->d_pop[h->cur], sizeof(double) * st.size(), hipMemcpyDeviceToHost, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++) {
                fr[Q * i + q] = st[(size_t)q * nl + i];
                fb[Q * i + q] = st[(size_t)(Q + q) * nl + i];
            }
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_get_fields(psx_lbm *h, double *rr, double *rb, double *cf) {
    PSX_CHECK_ARG(h, "psx_lbm_get_fields: NULL handle");
    PSX_GUARD_BEGIN
    std::vector<double> fr((size_t)Q * h->nl), fb((size_t)Q * h->nl);
    int rc = psx_lbm_get_state(h, fr.data(), fb.data());
    if (rc) return rc;
    for (int64_t i = 0; i < h->nl; i++) {
        double a = fr[Q * i], b = fb[Q * i];
        for (int q = 1; q < Q; q++) { a = a + fr[Q * i + q]; b = b + fb[Q * i + q]; }
        if (rr) rr[i] = a;
        if (rb) rb[i] = b;
        if (cf) cf[i] = (a - b) / (a + b);
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_fields(psx_lbm *h, double *u, double *color, double *pressure, int32_t *phase) {
    PSX_CHECK_ARG(h, "psx_lbm_fields: NULL handle");
    PSX_GUARD_BEGIN
    const int64_t nl = h->nl, nv = (int64_t)h->nx * h->ny * h->nz;
    // solid voxels (L1:2133-2136); fluid voxels are overwritten below
    for (int64_t a = 0; a < nv; a++) {
        if (u) u[3 * a] = u[3 * a + 1] = u[3 * a + 2] = 0.0;
        if (color) color[a] = 1.2;
        if (pressure) pressure[a] = 0.0;
        if (phase) phase[a] = 0;
    }
    if (nl == 0) return PSX_OK;
    if (h->fused) {  // the fused step keeps the colour field in LDS only
        int rc = launch_density(h, h->cur);
        if (rc) return rc;
    }
    std::vector<double> out((size_t)5 * nl);
    std::vector<int32_t> ph((size_t)nl);
    double *d_out = nullptr;
    int32_t *d_ph = nullptr;
    hipError_t e = hipMalloc((void **)&d_out, sizeof(double) * out.size());
    if (e == hipSuccess) e = hipMalloc((void **)&d_ph, sizeof(int32_t) * ph.size());
    if (e == hipSuccess) {
        const unsigned nb = (unsigned)psx::ceil_div(nl, 256);
        if (h->f32) {
            e = psx_lbm_fast::fields(h->c, (const float *)h->d_pop[h->cur], popx(h, h->cur), h->d_dst, h->d_vox,
                                     (const double *)h->d_phi, d_out, d_ph, h->stream);
        } else {
            StoreF64 st{(double *)h->d_pop[h->cur], nl};
            hipLaunchKernelGGL(k_fields<StoreF64>, dim3(nb), dim3(256), 0, h->stream, h->c, st, h->d_vox,
                               (const double *)h->d_phi, d_out, d_ph);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ph.data(), d_ph, sizeof(int32_t) * ph.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (d_out) (void)hipFree(d_out);
    if (d_ph) (void)hipFree(d_ph);
    PSX_HIP(e);
    for (int64_t i = 0; i < nl; i++) {
        const int32_t *ic = &h->icor[3 * h->ref_node(i)];
        const int64_t a = (int64_t)(ic[0] - 1) + (int64_t)h->nx * ((ic[1] - 1) + (int64_t)h->ny * (ic[2] - 1));
        if (u)
            for (int d = 0; d < 3; d++) u[3 * a + d] = out[(size_t)d * nl + i];
        if (color) color[a] = out[(size_t)3 * nl + i];
        if (pressure) pressure[a] = out[(size_t)4 * nl + i];
        if (phase) phase[a] = ph[i];
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_init_random(psx_lbm *h, double sat) {
    PSX_CHECK_ARG(h, "psx_lbm_init_random: NULL handle");
    PSX_GUARD_BEGIN
    // initdistr icon=0 (L1:2822-2841): Park-Miller draws in node order
    std::vector<double> fr((size_t)Q * h->nl), fb((size_t)Q * h->nl);
    double seed = 13.081984;
    for (int64_t i = 0; i < h->nl; i++) {
        seed = std::fmod(16807.0 * seed, 2147483647.0);
        double rn = seed / 2147483648.0;
        double cfield = rn <= sat ? 0.0 : 1.0;
        for (int q = 0; q < Q; q++) {
            fr[Q * i + q] = h->c.tr[q] * cfield * h->p.rr0;
            fb[Q * i + q] = h->c.tb[q] * (1.0 - cfield) * h->p.rb0;
        }
    }
    return psx_lbm_set_state(h, fr.data(), fb.data());
    PSX_GUARD_END
}

extern "C" int psx_lbm_step(psx_lbm *h, int64_t nsteps) {
    PSX_CHECK_ARG(h && nsteps >= 0, "psx_lbm_step: bad argument");
    for (int64_t s = 0; s < nsteps; s++) {
        int rc = launch_collide(h);
        if (rc) return rc;
        if (!h->fused) rc = launch_density(h, h->cur);
        if (rc) return rc;
    }
    return PSX_OK;
}

extern "C" int psx_lbm_sync(psx_lbm *h) {
    PSX_CHECK_ARG(h, "psx_lbm_sync: NULL handle");
    PSX_HIP(hipStreamSynchronize(h->stream));
    return PSX_OK;
}

extern "C" int psx_lbm_profile(psx_lbm *h, int64_t nsteps, double *ms) {
    PSX_CHECK_ARG(h && ms && nsteps > 0, "psx_lbm_profile: bad argument");
    PSX_GUARD_BEGIN
    psx::CallScope scope;  // destroys every event on every way out
    std::vector<hipEvent_t> ev(2 * nsteps + 1, nullptr);
    for (auto &e : ev) {
        PSX_HIP(hipEventCreate(&e));
        scope.event(e);
    }
    PSX_HIP(hipEventRecord(ev[0], h->stream));
    for (int64_t s = 0; s < nsteps; s++) {
        int rc = launch_collide(h);
        if (rc) return rc;
        PSX_HIP(hipEventRecord(ev[2 * s + 1], h->stream));
        if (!h->fused) rc = launch_density(h, h->cur);
        if (rc) return rc;
        PSX_HIP(hipEventRecord(ev[2 * s + 2], h->stream));
    }
    PSX_HIP(hipEventSynchronize(ev[2 * nsteps]));
    double tc = 0.0, td = 0.0;
    for (int64_t s = 0; s < nsteps; s++) {
        float a = 0.f, b = 0.f;
        PSX_HIP(hipEventElapsedTime(&a, ev[2 * s], ev[2 * s + 1]));
        PSX_HIP(hipEventElapsedTime(&b, ev[2 * s + 1], ev[2 * s + 2]));
        tc += a; td += b;
    }
    float tot = 0.f;
    PSX_HIP(hipEventElapsedTime(&tot, ev[0], ev[2 * nsteps]));
    ms[0] = tot;
    ms[1] = tc / (double)nsteps;
    ms[2] = td / (double)nsteps;
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_footprint(psx_lbm *h, int64_t *bytes) {
    PSX_CHECK_ARG(h && bytes, "psx_lbm_footprint: NULL argument");
    *bytes = (int64_t)Q * (h->f32 ? 2 * 4 + (h->f32x ? 1 : 0) : 2 * 8);
    return PSX_OK;
}

extern "C" int psx_lbm_check_state(psx_lbm *h, int64_t it1, int64_t itw, psx_lbm_diag *d) {
    PSX_CHECK_ARG(h && d, "psx_lbm_check_state: NULL argument");
    std::memset(d, 0, sizeof *d);
    double s[NDIAG] = {0};
    int64_t nb = psx::ceil_div(h->nl, 256);
    if (nb > 0) {
        if (h->fused) {  // the fused step keeps the colour field in LDS only
            int rc = launch_density(h, h->cur);
            if (rc) return rc;
        }
        if (h->f32) {
            PSX_HIP(psx_lbm_fast::diag(h->c, (const float *)h->d_pop[h->cur], popx(h, h->cur), h->d_dst, h->d_vox,
                                       (const double *)h->d_phi, h->d_partial, h->stream));
        } else {
            StoreF64 st{(double *)h->d_pop[h->cur], h->nl};
            hipLaunchKernelGGL(k_diag<StoreF64>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, st, h->d_vox,
                               (const double *)h->d_phi, h->d_partial);
        }
        PSX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_cols, dim3(NDIAG), dim3(256), 0, h->stream, h->d_partial, nb, NDIAG, h->d_red);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipMemcpyAsync(s, h->d_red, sizeof s, hipMemcpyDeviceToHost, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    }
    const psx_lbm_params &p = h->p;
    const double dti = 1.0 / p.dt;
    double svr = 1.0 / (3.0 * p.vir * p.dt + 0.5), svb = 1.0 / (3.0 * p.vib * p.dt + 0.5);
    double vir = dti * (1.0 / svr - 0.5) / 3.0, vib = dti * (1.0 / svb - 0.5) / 3.0;  // L1:3419-3420
    d->tmr = s[0]; d->tmb = s[1]; d->tm = s[2];
    d->Cpr = p.cr2 * s[4] / s[3];
    d->Cpb = h->c.cb2 * s[6] / s[5];
    d->isneg = s[7] > 0; d->negdens_r = s[8] > 0; d->negdens_b = s[9] > 0;
    double t0 = (double)((int32_t)(h->nx * h->ny * h->nz));
    for (int k = 0; k < 3; k++) {
        d->Qr[k] = s[13 + k] / t0; d->Qb[k] = s[16 + k] / t0; d->um[k] = s[19 + k] / t0;
        d->umr[k] = s[22 + k]; d->umb[k] = s[25 + k]; d->tfr[k] = s[28 + k]; d->tfb[k] = s[31 + k];
    }
    d->nr = s[34]; d->nb = s[35];
    int id = p.idir - 1;
    d->permr = d->Qr[id] * vir / p.bfr[id];
    d->permb = d->Qb[id] * vib / p.bfb[id];
    d->Re = p.lsize * s[12] / (s[1] * vir);
    double nl = (double)h->nl, sig = h->c.sig;
    d->outm[0] = (double)(it1 * itw);
    d->outm[1] = d->permr;
    d->outm[2] = d->permb;
    d->outm[3] = 0.03125 * d->Qr[id] * vir / sig;
    d->outm[4] = 0.03125 * d->Qb[id] * vib / sig;
    d->outm[5] = 0.03125 * d->umr[id] * vir * p.rr0 / (d->nr * sig);
    d->outm[6] = 0.03125 * d->umb[id] * vib * p.rb0 / (d->nb * sig);
    d->outm[7] = d->Cpr - d->Cpb;
    d->outm[8] = d->Cpr;
    d->outm[9] = d->Cpb;
    d->outm[10] = d->nr / nl;
    d->outm[11] = d->nb / nl;
    return PSX_OK;
}

// ============================================================ coupled ====
// L1's one-way coupled model: the solid lattice of the LBM's own ind (nodes
// at the corners of solid voxels, periodic in x, y and z; alf, bet from
// Em, nu, dl as L1:497-498) moves under the momentum the bounce-back links
// deliver each main-loop step (L1:3849-3879) by lsm_step (L1:1721-1903).
// The solid lives on an LSM handle without imposed strain; its state is
// [3][N] on the periodic node grid, mapped to L1's solid numbering (nodes
// z outer / x fastest, the order of the grid) on the host.
struct psx_cpl {
    psx_lbm *lbm = nullptr;
    psx_lsm *lsm = nullptr;
    int64_t ns = 0, N = 0;
    double alf = 0, bet = 0, rs = 1, dl = 1;
    std::vector<int32_t> grid;   // [ns] node-grid index of L1 solid node s
    std::vector<int32_t> mijks;  // L1 solid ids on the node grid (0 = none)
    CplLinks L{};
    int32_t *d_start = nullptr, *d_lnode = nullptr, *d_grid = nullptr;
    uint32_t *d_contrib = nullptr;
    int8_t *d_lq = nullptr;
    double *d_y = nullptr;
    double *d_af = nullptr;
    int32_t cbon[24][3], cang[72][6];
    double csb[24], cab[72];
    void *d_tab = nullptr;               // calc_ab tables (CubeTables, device)
    int8_t *d_raw = nullptr;             // ind codes as read (device)
    double *d_part = nullptr;            // calc_ab block partials
};

namespace {

struct FOutF64 {
    const double *p;
    int64_t nl;
    __device__ __forceinline__ double operator()(int64_t n, int qb) const {
        return p[(int64_t)qb * nl + n] + p[(int64_t)(Q + qb) * nl + n];
    }
};

// The 48 angle types of the D3Q19 lattice (pairs of arm directions, 1-based
// D3Q19 indices) and their plane normals and arm force directions
// anf(9,48): n, fj = n x xj, fk = -(n x xk), x = c/|c|^2, with the
// reference's constants 1/2, 1/sqrt(3), 1/(2 sqrt(3)) (XX:200-362).
void angle_table(int32_t nta[48][2], double anf[48][9]) {
    static const int r8[3][8] = {{2, 8, 4, 9, 3, 11, 5, 10}, {2, 12, 6, 13, 3, 15, 7, 14},
                                 {4, 16, 6, 17, 5, 19, 7, 18}};
    static const int r4[6][4] = {{12, 16, 13, 17}, {14, 18, 15, 19}, {8, 12, 10, 14},
                                 {9, 13, 11, 15},  {8, 16, 9, 18},   {10, 17, 11, 19}};
    int t = 0;
    for (auto &ring : r8)
        for (int m = 0; m < 8; m++, t++) { nta[t][0] = ring[m]; nta[t][1] = ring[(m + 1) % 8]; }
    for (auto &ring : r4)
        for (int m = 0; m < 4; m++, t++) { nta[t][0] = ring[m]; nta[t][1] = ring[(m + 1) % 4]; }
    const double s3 = 1.0 / std::sqrt(3.0), s23 = 1.0 / (2.0 * std::sqrt(3.0));
    for (t = 0; t < 48; t++) {
        const int *cj = H_IV[nta[t][0] - 1], *ck = H_IV[nta[t][1] - 1];
        const int sg = (t >= 8 && t < 16) ? -1 : 1;  // the reference's orientation of the xz ring
        const int n[3] = {sg * (cj[1] * ck[2] - cj[2] * ck[1]), sg * (cj[2] * ck[0] - cj[0] * ck[2]),
                          sg * (cj[0] * ck[1] - cj[1] * ck[0])};
        const int n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        const int gj[3] = {n[1] * cj[2] - n[2] * cj[1], n[2] * cj[0] - n[0] * cj[2], n[0] * cj[1] - n[1] * cj[0]};
        const int gk[3] = {n[2] * ck[1] - n[1] * ck[2], n[0] * ck[2] - n[2] * ck[0], n[1] * ck[0] - n[0] * ck[1]};
        const int cj2 = cj[0] * cj[0] + cj[1] * cj[1] + cj[2] * cj[2];
        const int ck2 = ck[0] * ck[0] + ck[1] * ck[1] + ck[2] * ck[2];
        auto comp = [&](int g, int c2) -> double {  // g / (|n| |c|^2)
            if (g == 0) return 0.0;
            if (n2 == 1) return c2 == 1 ? (double)g : 0.5 * (double)g;
            return (g > 0 ? 1.0 : -1.0) * (std::abs(g) == 2 ? s3 : s23);
        };
        for (int d = 0; d < 3; d++) {
            anf[t][d] = n[d] == 0 ? 0.0 : (n2 == 1 ? (double)n[d] : (n[d] > 0 ? s3 : -s3));
            anf[t][3 + d] = comp(gj[d], cj2);
            anf[t][6 + d] = comp(gk[d], ck2);
        }
    }
}

// Tables of calc_ab's per-cube loops (L1:4101-4113, 996-1014).
struct CubeTables {
    int32_t bon[24][3], ang[72][6];
    double sb[24], ab[72];  // cube_sbond * alf, cube_abond * bet
    double anf[48][9];
    double bn[9][3], vb[9][3];
    double vn[Q][3], vx[Q][3], vv[Q][3];
};

constexpr int NAB = 20;  // per-cube sums: sigmap (9), strain / 4 (9), cubes, pore voxels

// calc_ab's cube loop (L1:4170-4340) with one thread per voxel: every solid
// voxel evaluates its cube's stress (sigmap; sigmam is its exact negative)
// and strain in the reference's operation order; the block sums go to
// partial[block][NAB].  r is the displacement on the periodic node grid,
// [3][N], node (x,y,z) at x + nx (y + ny z).
__global__ __launch_bounds__(256) void k_calc_ab_cubes(const CubeTables *__restrict__ T,
                                                       const int8_t *__restrict__ raw, int nx, int ny, int nz,
                                                       int iipore, const double *__restrict__ r, int64_t N,
                                                       double *__restrict__ partial) {
    const int64_t nv = (int64_t)nx * ny * nz;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[NAB];
#pragma unroll
    for (int k = 0; k < NAB; k++) acc[k] = 0.0;
    if (v < nv) {
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((int64_t)nx * ny));
        const int code = raw[v];
        if (code == iipore) acc[19] = 1.0;
        if (code == 0) {
            acc[18] = 1.0;
            int64_t g[8];
            for (int l = 0; l < 8; l++) {
                const int x = (i + (l & 1)) % nx, y = (j + ((l >> 1) & 1)) % ny, z = (k + (l >> 2)) % nz;
                g[l] = x + (int64_t)nx * (y + (int64_t)ny * z);
            }
            auto R = [&](int l, int d) { return r[(int64_t)d * N + g[l]]; };
            double sp[3][3] = {{0.0}}, st[3][3] = {{0.0}};  // [column][row]
            for (int l = 0; l < 8; l++) {
                const double sv[3] = {(l & 1) ? 0.5 : -0.5, ((l >> 1) & 1) ? 0.5 : -0.5, (l >> 2) ? 0.5 : -0.5};
                const double up[3] = {R(l, 0), R(l, 1), R(l, 2)};
                for (int m = 0; m < 3; m++)
                    for (int n = 0; n < 3; n++) st[n][m] = st[n][m] + sv[m] * up[n] + sv[n] * up[m];
            }
            for (int l = 0; l < 24; l++) {
                const int p1 = T->bon[l][0] - 1, p2 = T->bon[l][1] - 1, b = T->bon[l][2] - 1;
                double rt[3], fl[3];
                for (int d = 0; d < 3; d++) rt[d] = R(p2, d) - R(p1, d) + T->vb[b][d];
                const double dot = rt[0] * T->bn[b][0] + rt[1] * T->bn[b][1] + rt[2] * T->bn[b][2];
                for (int d = 0; d < 3; d++) fl[d] = T->sb[l] * (dot * T->bn[b][d] - T->vb[b][d]);
                for (int cd = 0; cd < 3; cd++) {
                    if (T->bn[b][cd] > 0)
                        for (int d = 0; d < 3; d++) sp[cd][d] = sp[cd][d] + fl[d];
                    if (T->bn[b][cd] < 0)
                        for (int d = 0; d < 3; d++) sp[cd][d] = sp[cd][d] - fl[d];
                }
            }
            for (int l = 0; l < 72; l++) {
                const int32_t *e = T->ang[l];
                const int pi = e[0] - 1, pj = e[1] - 1, pk = e[2] - 1, qj = e[3] - 1, qk = e[4] - 1;
                const double *an = T->anf[e[5] - 1];
                const double *xrj = T->vx[qj], *xrk = T->vx[qk];
                double rtj[3], rtk[3];
                for (int d = 0; d < 3; d++) {
                    rtj[d] = R(pj, d) - R(pi, d) + T->vv[qj][d];
                    rtk[d] = R(pk, d) - R(pi, d) + T->vv[qk][d];
                }
                const double dfi = ((xrk[1] * rtk[2] - xrk[2] * rtk[1] - xrj[1] * rtj[2] + xrj[2] * rtj[1]) * an[0] -
                                    (xrk[0] * rtk[2] - xrk[2] * rtk[0] - xrj[0] * rtj[2] + xrj[2] * rtj[0]) * an[1] +
                                    (xrk[0] * rtk[1] - xrk[1] * rtk[0] - xrj[0] * rtj[1] + xrj[1] * rtj[0]) * an[2]) *
                                   T->ab[l];
                for (int cd = 0; cd < 3; cd++) {
                    if (T->vn[qj][cd] > 0)
                        for (int d = 0; d < 3; d++) sp[cd][d] = sp[cd][d] - dfi * an[3 + d];
                    if (T->vn[qj][cd] < 0)
                        for (int d = 0; d < 3; d++) sp[cd][d] = sp[cd][d] + dfi * an[3 + d];
                }
                for (int cd = 0; cd < 3; cd++) {
                    if (T->vn[qk][cd] > 0)
                        for (int d = 0; d < 3; d++) sp[cd][d] = sp[cd][d] - dfi * an[6 + d];
                    if (T->vn[qk][cd] < 0)
                        for (int d = 0; d < 3; d++) sp[cd][d] = sp[cd][d] + dfi * an[6 + d];
                }
            }
            for (int c = 0; c < 3; c++)
                for (int d = 0; d < 3; d++) {
                    acc[d + 3 * c] = sp[c][d];
                    acc[9 + d + 3 * c] = st[c][d] * 0.25;
                }
        }
    }
    __shared__ double sh[4][NAB];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < NAB; k++) {
        const double t = psx::wave_sum(acc[k]);
        if (lane == 0) sh[w][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < NAB) {
        const int k = threadIdx.x;
        partial[(int64_t)blockIdx.x * NAB + k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}

constexpr int MAXPORES = 9;  // pore family codes are single digits

// calc_ab's beta loop (L1:4381-4445): every face between a voxel of pore
// family f and a solid voxel adds -(u_face . c) to family f, u_face the mean
// displacement of the face's 4 nodes; block sums per family.
__global__ __launch_bounds__(256) void k_calc_ab_faces(const int8_t *__restrict__ raw, int nx, int ny, int nz,
                                                       int npores, const double *__restrict__ r, int64_t N,
                                                       double *__restrict__ partial) {
    const int64_t nv = (int64_t)nx * ny * nz;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[MAXPORES];
#pragma unroll
    for (int k = 0; k < MAXPORES; k++) acc[k] = 0.0;
    if (v < nv) {
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((int64_t)nx * ny));
        const int code = raw[v];
        if (code >= 1 && code <= npores) {
            double b = 0.0;
            for (int l = 1; l <= 6; l++) {
                const int c0 = H_IV[l][0], c1 = H_IV[l][1], c2 = H_IV[l][2];
                const int xn = (i + c0 + nx) % nx, yn = (j + c1 + ny) % ny, zn = (k + c2 + nz) % nz;
                if (raw[xn + (int64_t)nx * (yn + (int64_t)ny * zn)] != 0) continue;
                double up[3] = {0.0, 0.0, 0.0};
                int64_t g[4];
                for (int m = 0; m < 4; m++) {
                    const int o1 = m & 1, o2 = m >> 1;
                    int off[3];
                    if (c0 != 0) { off[0] = (c0 + 1) / 2; off[1] = o1; off[2] = o2; }
                    else if (c1 != 0) { off[0] = o1; off[1] = (c1 + 1) / 2; off[2] = o2; }
                    else { off[0] = o1; off[1] = o2; off[2] = (c2 + 1) / 2; }
                    g[m] = (i + off[0]) % nx + (int64_t)nx * ((j + off[1]) % ny + (int64_t)ny * ((k + off[2]) % nz));
                }
                for (int d = 0; d < 3; d++)
                    up[d] = (r[d * N + g[0]] + r[d * N + g[1]] + r[d * N + g[2]] + r[d * N + g[3]]) * 0.25;
                b = b - up[0] * (double)c0 - up[1] * (double)c1 - up[2] * (double)c2;
            }
            for (int f = 0; f < MAXPORES; f++)
                if (f == code - 1) acc[f] = b;
        }
    }
    __shared__ double sh[4][MAXPORES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int f = 0; f < MAXPORES; f++) {
        const double t = psx::wave_sum(acc[f]);
        if (lane == 0) sh[w][f] = t;
    }
    __syncthreads();
    if (threadIdx.x < MAXPORES) {
        const int f = threadIdx.x;
        partial[(int64_t)blockIdx.x * MAXPORES + f] = ((sh[0][f] + sh[1][f]) + sh[2][f]) + sh[3][f];
    }
}

}  // namespace

extern "C" int psx_cpl_create(psx_lbm *lbm, double Em, double nu, double dl, double rs, const int32_t *cube_bon,
                              const int32_t *cube_ang, const double *cube_sbond, const double *cube_abond,
                              psx_cpl **out) {
    PSX_CHECK_ARG(lbm && out && cube_bon && cube_ang && cube_sbond && cube_abond,
                  "psx_cpl_create: NULL argument (the cube tables are required)");
    PSX_CHECK_ARG(rs > 0.0 && dl > 0.0, "psx_cpl_create: rs and dl must be > 0");
    PSX_CHECK_ARG(5.0 - 10.0 * nu != 0.0 && 2 * nu * nu + nu - 1 != 0.0, "psx_cpl_create: nu gives no springs");
    *out = nullptr;
    PSX_GUARD_BEGIN
    psx_cpl *c = new psx_cpl();
    psx::CreateGuard<psx_cpl> guard(c, psx_cpl_destroy);
    c->lbm = lbm;
    c->rs = rs;
    c->dl = dl;
    c->alf = dl * Em / (5.0 - 10.0 * nu);                                      // L1:497
    c->bet = dl * dl * dl * Em * (4 * nu - 1) / (20.0 * (2 * nu * nu + nu - 1));  // L1:498
    std::memcpy(c->cbon, cube_bon, sizeof c->cbon);
    std::memcpy(c->cang, cube_ang, sizeof c->cang);
    std::memcpy(c->csb, cube_sbond, sizeof c->csb);
    std::memcpy(c->cab, cube_abond, sizeof c->cab);
    const int nx = lbm->nx, ny = lbm->ny, nz = lbm->nz;
    const int64_t nv = (int64_t)nx * ny * nz;
    // the solid lattice: voxels with ind == 0 are solid, every other code is
    // pore for the springs (L1:501-510)
    std::vector<int32_t> ind(nv);
    for (int64_t a = 0; a < nv; a++) ind[a] = lbm->ind_raw[a] == 0 ? 0 : 1;
    psx_lsm_params lp{};
    lp.alf1 = lp.alf2 = c->alf;
    lp.bet1 = lp.bet2 = c->bet;
    lp.rs = rs;
    lp.dt = lbm->p.dt;
    lp.gstrain = 0.0;
    lp.precision = PSX_PREC_F64;
    int rc = psx_lsm_create(&lp, nx, ny, nz, ind.data(), nullptr, cube_bon, cube_ang, cube_sbond, cube_abond, &c->lsm);
    if (rc) return rc;
    c->N = psx_lsm_cpl::grid_nodes(c->lsm);
    coupling_nodes(lbm, c->mijks);
    for (int64_t g = 0; g < nv; g++)
        if (c->mijks[g] > 0) c->grid.push_back((int32_t)g);
    c->ns = (int64_t)c->grid.size();
    // links: ipb2 rows, then ipb4 rows (the reference's loop order)
    const int64_t n2 = lbm->npb2, n4 = lbm->npb4;
    std::vector<int32_t> ipb2(4 * (size_t)n2), ipb4(6 * (size_t)n4);
    rc = psx_lbm_tables(lbm, nullptr, nullptr, ipb2.data(), ipb4.data());
    if (rc) return rc;
    std::vector<int32_t> inv;
    if (!lbm->perm.empty()) {
        inv.resize(lbm->perm.size());
        for (size_t i = 0; i < lbm->perm.size(); i++) inv[(size_t)lbm->perm[i]] = (int32_t)i;
    }
    const int64_t nlink = n2 + n4;
    if (nlink >= ((int64_t)1 << CPL_QSHIFT)) {
        psx::set_error("psx_cpl_create: %lld bounce-back links exceed the coupling table", (long long)nlink);
        return PSX_EINVAL;
    }
    std::vector<int32_t> lnode((size_t)std::max<int64_t>(nlink, 1)), cnt((size_t)c->ns + 1, 0);
    std::vector<int8_t> lq((size_t)std::max<int64_t>(nlink, 1));
    auto row = [&](int64_t li) -> const int32_t * { return li < n2 ? &ipb2[4 * li] : &ipb4[6 * (li - n2)]; };
    for (int64_t li = 0; li < nlink; li++) {
        const int32_t *e = row(li);
        const int64_t n = e[0] - 1;
        lnode[(size_t)li] = (int32_t)(inv.empty() ? n : inv[(size_t)n]);
        lq[(size_t)li] = (int8_t)(e[1] - 1);
        for (int m = 0; m < (li < n2 ? 2 : 4); m++)
            if (e[2 + m] > 0) cnt[(size_t)e[2 + m] - 1]++;  // id 0: a wall voxel of another pore family
    }
    // surface solid nodes (those with links), their contribution ranges; a
    // node's links sit next to it in the fluid numbering (x outer, z
    // fastest), so the surface list follows that order too and neighbouring
    // threads read neighbouring per-link terms
    std::vector<int32_t> order;
    for (int64_t s = 0; s < c->ns; s++)
        if (cnt[(size_t)s] > 0) order.push_back((int32_t)s);
    auto fluid_key = [&](int32_t s) {
        const int64_t g = c->grid[(size_t)s];
        const int64_t x = g % nx, y = (g / nx) % ny, z = g / ((int64_t)nx * ny);
        return (x * ny + y) * (int64_t)nz + z;
    };
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return fluid_key(a) < fluid_key(b); });
    std::vector<int32_t> sidx((size_t)c->ns, -1), sgrid, start(1, 0);
    for (int32_t s : order) {
        sidx[(size_t)s] = (int32_t)sgrid.size();
        sgrid.push_back(c->grid[(size_t)s]);
        start.push_back(start.back() + cnt[(size_t)s]);
    }
    const int64_t nsurf = (int64_t)sgrid.size();
    std::vector<uint32_t> contrib((size_t)std::max<int32_t>(start.back(), 1));
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int64_t li = 0; li < nlink; li++) {
        const int32_t *e = row(li);
        for (int m = 0; m < (li < n2 ? 2 : 4); m++)
            if (e[2 + m] > 0)
                contrib[(size_t)fill[(size_t)sidx[(size_t)e[2 + m] - 1]]++] =
                    (uint32_t)li | ((uint32_t)(e[1] - 1) << CPL_QSHIFT);
    }
    PSX_HIP(hipMalloc(&c->d_start, sizeof(int32_t) * start.size()));
    PSX_HIP(hipMalloc(&c->d_contrib, sizeof(uint32_t) * contrib.size()));
    PSX_HIP(hipMalloc(&c->d_lnode, sizeof(int32_t) * lnode.size()));
    PSX_HIP(hipMalloc(&c->d_lq, lq.size()));
    PSX_HIP(hipMalloc(&c->d_y, sizeof(double) * lnode.size()));
    PSX_HIP(hipMalloc(&c->d_grid, sizeof(int32_t) * std::max<size_t>(1, sgrid.size())));
    PSX_HIP(hipMalloc(&c->d_af, sizeof(double) * 3 * (size_t)c->N));
    hipStream_t st = lbm->stream;
    PSX_HIP(hipMemcpyAsync(c->d_start, start.data(), sizeof(int32_t) * start.size(), hipMemcpyHostToDevice, st));
    PSX_HIP(hipMemcpyAsync(c->d_contrib, contrib.data(), sizeof(uint32_t) * contrib.size(), hipMemcpyHostToDevice, st));
    PSX_HIP(hipMemcpyAsync(c->d_lnode, lnode.data(), sizeof(int32_t) * lnode.size(), hipMemcpyHostToDevice, st));
    PSX_HIP(hipMemcpyAsync(c->d_lq, lq.data(), lq.size(), hipMemcpyHostToDevice, st));
    if (nsurf > 0)
        PSX_HIP(hipMemcpyAsync(c->d_grid, sgrid.data(), sizeof(int32_t) * sgrid.size(), hipMemcpyHostToDevice, st));
    // af stays zero at solid nodes without links
    PSX_HIP(hipMemsetAsync(c->d_af, 0, sizeof(double) * 3 * (size_t)c->N, st));
    PSX_HIP(hipStreamSynchronize(st));
    CplLinks &L = c->L;
    L.lnode = c->d_lnode; L.lq = c->d_lq; L.y = c->d_y;
    L.start = c->d_start; L.contrib = c->d_contrib; L.sgrid = c->d_grid;
    L.nlink = (int32_t)nlink;
    L.nsurf = (int32_t)nsurf;
    L.nedge = (int32_t)n2;
    L.N = c->N;
    const psx_lbm_params &p = lbm->p;
    const double dt2 = p.dt * p.dt, dti = 1.0 / p.dt;
    for (int q = 0; q < Q; q++) {
        L.fq[q] = q == 0 ? p.rr0 * 0.0 : p.rr0 * (q < 7 ? p.cr2 * (1.0 / 6.0) * dt2 : p.cr2 * (1.0 / 12.0) * dt2);
        for (int d = 0; d < 3; d++) L.cv[q][d] = (double)H_IV[q][d];
    }
    L.rsdti2 = (1.0 / rs) * dti * dti;
    {  // calc_ab tables and the raw voxel codes
        CubeTables T{};
        std::memcpy(T.bon, c->cbon, sizeof T.bon);
        std::memcpy(T.ang, c->cang, sizeof T.ang);
        for (int l = 0; l < 24; l++) T.sb[l] = c->csb[l] * c->alf;
        for (int l = 0; l < 72; l++) T.ab[l] = c->cab[l] * c->bet;
        int32_t nta[48][2];
        angle_table(nta, T.anf);
        static const int IB[9][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1},  {1, 1, 0}, {1, -1, 0},
                                     {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
        for (int q = 0; q < Q; q++) {
            double s2 = 0.0;
            for (int d = 0; d < 3; d++) { T.vv[q][d] = H_IV[q][d]; s2 += T.vv[q][d] * T.vv[q][d]; }
            const double nr = std::sqrt(s2);
            for (int d = 0; d < 3; d++) { T.vn[q][d] = q ? T.vv[q][d] / nr : 0.0; T.vx[q][d] = q ? T.vv[q][d] / s2 : 0.0; }
        }
        for (int b = 0; b < 9; b++) {
            double s2 = 0.0;
            for (int d = 0; d < 3; d++) { T.vb[b][d] = IB[b][d]; s2 += T.vb[b][d] * T.vb[b][d]; }
            const double nr = std::sqrt(s2);
            for (int d = 0; d < 3; d++) T.bn[b][d] = T.vb[b][d] / nr;
        }
        std::vector<int8_t> raw8((size_t)nv);
        for (int64_t a = 0; a < nv; a++) raw8[(size_t)a] = (int8_t)lbm->ind_raw[(size_t)a];
        const int64_t nb = psx::ceil_div(nv, 256);
        PSX_HIP(hipMalloc(&c->d_tab, sizeof(CubeTables)));  // holds T
        PSX_HIP(hipMalloc(&c->d_raw, (size_t)nv));
        PSX_HIP(hipMalloc(&c->d_part, sizeof(double) * (size_t)nb * NAB + sizeof(double) * NAB));
        PSX_HIP(hipMemcpyAsync(c->d_tab, &T, sizeof T, hipMemcpyHostToDevice, st));
        PSX_HIP(hipMemcpyAsync(c->d_raw, raw8.data(), (size_t)nv, hipMemcpyHostToDevice, st));
        PSX_HIP(hipStreamSynchronize(st));
    }
    rc = psx_lsm_cpl::reset(c->lsm);
    if (rc) return rc;
    *out = guard.release();
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" void psx_cpl_destroy(psx_cpl *c) {
    if (!c) return;
    if (c->lbm && c->lbm->stream) (void)hipStreamSynchronize(c->lbm->stream);
    void *bufs[] = {c->d_start, c->d_contrib, c->d_lnode, c->d_lq, c->d_y, c->d_grid, c->d_af, c->d_tab, c->d_raw,
                    c->d_part};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->lsm) psx_lsm_destroy(c->lsm);
    delete c;
}

extern "C" int psx_cpl_counts(psx_cpl *c, int64_t *counts) {
    PSX_CHECK_ARG(c && counts, "psx_cpl_counts: NULL argument");
    counts[0] = c->ns;
    counts[1] = c->L.nedge;
    counts[2] = c->lbm->npb2 + c->lbm->npb4;
    counts[3] = c->N;
    return PSX_OK;
}

extern "C" int psx_cpl_nodes(psx_cpl *c, int32_t *icors, int32_t *mijks) {
    PSX_CHECK_ARG(c, "psx_cpl_nodes: NULL handle");
    const int nx = c->lbm->nx, ny = c->lbm->ny;
    for (int64_t s = 0; s < c->ns; s++) {
        const int64_t g = c->grid[(size_t)s];
        if (icors) {
            icors[3 * s] = (int32_t)(g % nx) + 1;
            icors[3 * s + 1] = (int32_t)((g / nx) % ny) + 1;
            icors[3 * s + 2] = (int32_t)(g / ((int64_t)nx * ny)) + 1;
        }
    }
    if (mijks) std::memcpy(mijks, c->mijks.data(), sizeof(int32_t) * c->mijks.size());
    return PSX_OK;
}

extern "C" int psx_cpl_step(psx_cpl *c, int64_t nsteps) {
    PSX_CHECK_ARG(c && nsteps >= 0, "psx_cpl_step: bad argument");
    psx_lbm *h = c->lbm;
    for (int64_t s = 0; s < nsteps; s++) {
        int rc = psx_lbm_step(h, 1);
        if (rc) return rc;
        if (c->L.nsurf > 0) {
            if (h->f32) {
                PSX_HIP(psx_lbm_fast::link_force(h->c, (const float *)h->d_pop[h->cur], c->L, h->stream));
            } else {
                FOutF64 f{(const double *)h->d_pop[h->cur], h->nl};
                hipLaunchKernelGGL(k_link_force<FOutF64>, dim3((unsigned)psx::ceil_div(c->L.nlink, 256)), dim3(256), 0,
                                   h->stream, c->L, f);
            }
            hipLaunchKernelGGL(k_node_force, dim3((unsigned)psx::ceil_div(c->L.nsurf, 256)), dim3(256), 0, h->stream,
                               c->L, c->d_af);
            PSX_HIP(hipGetLastError());
        }
        rc = psx_lsm_cpl::step(c->lsm, c->d_af, 0.25, h->stream);  // damping of L1:1888
        if (rc) return rc;
    }
    return PSX_OK;
}

extern "C" int psx_cpl_get_state(psx_cpl *c, double *r, double *v, double *a, double *af) {
    PSX_CHECK_ARG(c, "psx_cpl_get_state: NULL handle");
    PSX_GUARD_BEGIN
    double *dev[4];
    psx_lsm_cpl::state(c->lsm, &dev[0], &dev[1], &dev[2]);
    dev[3] = c->d_af;
    double *outp[4] = {r, v, a, af};
    std::vector<double> buf(3 * (size_t)c->N);
    for (int k = 0; k < 4; k++) {
        if (!outp[k]) continue;
        PSX_HIP(hipMemcpyAsync(buf.data(), dev[k], sizeof(double) * buf.size(), hipMemcpyDeviceToHost, c->lbm->stream));
        PSX_HIP(hipStreamSynchronize(c->lbm->stream));
        for (int64_t s = 0; s < c->ns; s++)
            for (int d = 0; d < 3; d++) outp[k][3 * s + d] = buf[(size_t)d * c->N + c->grid[(size_t)s]];
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_cpl_set_state(psx_cpl *c, const double *r, const double *v, const double *a) {
    PSX_CHECK_ARG(c && r && v && a, "psx_cpl_set_state: NULL argument");
    PSX_GUARD_BEGIN
    double *dev[3];
    psx_lsm_cpl::state(c->lsm, &dev[0], &dev[1], &dev[2]);
    const double *inp[3] = {r, v, a};
    std::vector<double> buf(3 * (size_t)c->N);
    for (int k = 0; k < 3; k++) {
        std::fill(buf.begin(), buf.end(), 0.0);
        for (int64_t s = 0; s < c->ns; s++)
            for (int d = 0; d < 3; d++) buf[(size_t)d * c->N + c->grid[(size_t)s]] = inp[k][3 * s + d];
        PSX_HIP(hipMemcpyAsync(dev[k], buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice, c->lbm->stream));
        PSX_HIP(hipStreamSynchronize(c->lbm->stream));
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_cpl_calc_ab(psx_cpl *c, int32_t npores, double pres, double *out) {
    PSX_CHECK_ARG(c && out && npores >= 0 && npores <= MAXPORES, "psx_cpl_calc_ab: bad argument (npores 0..9)");
    PSX_GUARD_BEGIN
    psx_lbm *h = c->lbm;
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const int64_t nv = (int64_t)nx * ny * nz, nb = psx::ceil_div(nv, 256);
    if (!(pres > 0.0)) pres = (1.0 - 0.99) * h->p.rr0 * h->p.cr2;  // L1:3987
    double *r, *v, *a;
    psx_lsm_cpl::state(c->lsm, &r, &v, &a);
    double *red = c->d_part + nb * NAB;
    double sums[NAB] = {0.0}, fam[MAXPORES] = {0.0};
    hipLaunchKernelGGL(k_calc_ab_cubes, dim3((unsigned)nb), dim3(256), 0, h->stream,
                       (const CubeTables *)c->d_tab, c->d_raw, nx, ny, nz,
                       (int)h->p.iipore, r, c->N, c->d_part);
    hipLaunchKernelGGL(k_reduce_cols, dim3(NAB), dim3(256), 0, h->stream, c->d_part, nb, NAB, red);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipMemcpyAsync(sums, red, sizeof sums, hipMemcpyDeviceToHost, h->stream));
    PSX_HIP(hipStreamSynchronize(h->stream));
    if (npores > 0) {
        hipLaunchKernelGGL(k_calc_ab_faces, dim3((unsigned)nb), dim3(256), 0, h->stream, c->d_raw, nx, ny, nz, npores,
                           r, c->N, c->d_part);
        hipLaunchKernelGGL(k_reduce_cols, dim3(MAXPORES), dim3(256), 0, h->stream, c->d_part, nb, MAXPORES, red);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipMemcpyAsync(fam, red, sizeof fam, hipMemcpyDeviceToHost, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    }
    const double vcell = (double)((int32_t)(nx * ny * nz));
    const double epsi = sums[19] / vcell;
    for (int k = 0; k < 9; k++) {
        const int row = k % 3, col = k / 3;
        out[k] = sums[k];
        out[9 + k] = -sums[k];  // sigmam mirrors sigmap term by term
        out[18 + k] = sums[9 + k];
        out[28 + k] = -(row == col ? 1.0 : 0.0) * epsi + sums[k] / vcell / pres;  // L1:4378
    }
    out[27] = epsi;
    out[37] = sums[18];
    for (int f = 0; f < npores; f++) out[38 + f] = fam[f] / vcell / pres;
    return PSX_OK;
    PSX_GUARD_END
}
