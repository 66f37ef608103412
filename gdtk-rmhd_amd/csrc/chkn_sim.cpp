// chkn_sim.cpp -- host runtime of libchkn: job set-up, geometry pre-pass, device memory,
// step orchestration (segments of steps between CFL updates, optionally replayed from a
// captured hipGraph) and the C-ABI of include/chkn.h.
//
// The march loop restates march_in_time_using_cpu_only (simulate.cu:182-342).  The time step
// only changes every cfl_count steps, so the host knows dt for a whole segment of steps:
// it writes dt into the device control block, issues the segment's 3*K stage launches
// without synchronising, and reads the control block back once at the end of the segment
// (bad-cell counts per stage and, when the segment ends on a CFL step, the minimum signal
// time reduced inside the last stage-3 launch).
//
// A handle may hold several instances: independent jobs, each with its own block array,
// inflow states, time, time step, CFL and plot schedules.  All instances advance in
// lockstep, one launch per stage for every block of every instance; an instance that has
// reached its max_step or max_time skips the remaining steps (its run flag is 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/chkn.h"
#include "chkn_geometry.h"
#include "chkn_internal.h"
#include "chkn_io.h"
#include "chkn_physics.h"

using namespace chkn;

#include "chkn_runtime.h"

namespace chkn {
namespace rt {

thread_local std::string g_create_error;
int g_verbosity = 0;  // Config::verbosity of the reference (config.cu:312-344)

int set_error(chkn_sim* s, int code, const std::string& msg)
{
    if (s) s->err = msg;
    else g_create_error = msg;
    return code;
}

real schedule_get(const std::vector<real>& ts, const std::vector<real>& vs, real t)
{
    // config.cu:263-269
    int i = (int)ts.size() - 1;
    while (i > 0 && t < ts[i]) i--;
    return vs[i];
}

std::vector<real> to_real(const double* v, size_t n)
{
    std::vector<real> r(n);
    for (size_t i = 0; i < n; ++i) r[i] = (real)v[i];
    return r;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// hipMalloc counted in the handle's set-up total (reported by chkn_device_bytes).
void dalloc(chkn_sim* s, void** p, size_t bytes)
{
    HIPCHK(hipMalloc(p, bytes));
    s->device_bytes += (long long)bytes;
}

void free_device(chkn_sim* s)
{
    free_grad(s);
    split_free(s);
    for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
    for (auto& b : s->blocks) {
        dfree(b.W); dfree(b.U); dfree(b.L); dfree(b.gi); dfree(b.gj); dfree(b.gk);
        dfree(b.volinv); dfree(b.len); dfree(b.vtx);
        b.W = b.U = b.L = b.gi = b.gj = b.gk = b.volinv = b.len = b.vtx = nullptr;
    }
    dfree(s->d_blocks); dfree(s->d_tiles); dfree(s->d_inflow); dfree(s->d_ctl);
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    s->d_blocks = nullptr; s->d_tiles = nullptr; s->d_inflow = nullptr; s->d_ctl = nullptr;
    s->h_ctl = nullptr; s->ev0 = s->ev1 = nullptr; s->stream = s->own_stream = nullptr;
}

void upload_blocks(chkn_sim* s)
{
    HIPCHK(hipMemcpy(s->d_blocks, s->bdev.data(), sizeof(BlockDev) * s->bdev.size(),
                     hipMemcpyHostToDevice));
}

// Geometry pre-pass of one block (block.cu:633-652) into device-layout host arrays.
// fpos, when given, receives the face centres [3][nft] (i-, then j-, then k-faces).
void block_geometry(HostBlock& b, const real* vtx, std::vector<real>& gi, std::vector<real>& gj,
                    std::vector<real>& gk, std::vector<real>& volinv, std::vector<real>& len,
                    std::vector<real>* fpos)
{
    const int nic = b.nic, njc = b.njc, nkc = b.nkc;
    auto V = [&](int i, int j, int k) {
        size_t q = (size_t)k * (nic + 1) * (njc + 1) + (size_t)j * (nic + 1) + i;
        return geom::V3{vtx[3 * q], vtx[3 * q + 1], vtx[3 * q + 2]};
    };
    const size_t n = (size_t)b.n;
    b.pos.assign(3 * n, 0.0);
    b.vol.assign(n, 0.0);
    volinv.assign(n, 0.0);
    len.assign(3 * n, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i < nic; ++i) {
                geom::V3 p[8] = {V(i, j, k),         V(i + 1, j, k),     V(i + 1, j + 1, k),
                                 V(i, j + 1, k),     V(i, j, k + 1),     V(i + 1, j, k + 1),
                                 V(i + 1, j + 1, k + 1), V(i, j + 1, k + 1)};
                geom::Hex h = geom::hex(p);
                size_t c = (size_t)i + (size_t)nic * (j + (size_t)njc * k);
                b.pos[c] = h.c.x; b.pos[n + c] = h.c.y; b.pos[2 * n + c] = h.c.z;
                b.vol[c] = h.vol;
                volinv[c] = real(1) / h.vol;  // cell.cu:190
                len[c] = h.iL; len[n + c] = h.jL; len[2 * n + c] = h.kL;
            }
    auto put = [](std::vector<real>& g, size_t nf, size_t f, const geom::Quad& q) {
        g[f] = q.area;
        g[nf + f] = q.n.x; g[2 * nf + f] = q.n.y; g[3 * nf + f] = q.n.z;
        g[4 * nf + f] = q.t1.x; g[5 * nf + f] = q.t1.y; g[6 * nf + f] = q.t1.z;
        g[7 * nf + f] = q.t2.x; g[8 * nf + f] = q.t2.y; g[9 * nf + f] = q.t2.z;
    };
    size_t nfi = (size_t)(nic + 1) * njc * nkc, nfj = (size_t)nic * (njc + 1) * nkc,
           nfk = (size_t)nic * njc * (nkc + 1);
    const size_t nft = nfi + nfj + nfk;
    if (fpos) fpos->assign(3 * nft, 0.0);
    auto centre = [&](size_t ff, const geom::Quad& q) {
        if (!fpos) return;
        (*fpos)[ff] = q.c.x; (*fpos)[nft + ff] = q.c.y; (*fpos)[2 * nft + ff] = q.c.z;
    };
    gi.assign(kFrameN * nfi, 0.0);
    gj.assign(kFrameN * nfj, 0.0);
    gk.assign(kFrameN * nfk, 0.0);
    // Face vertex orders: i-faces block.cu:267-270, j-faces :368-371, k-faces :469-472.
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i <= nic; ++i) {
                const size_t f = (size_t)i + (size_t)(nic + 1) * (j + (size_t)njc * k);
                const geom::Quad q = geom::quad(V(i, j, k), V(i, j + 1, k), V(i, j + 1, k + 1), V(i, j, k + 1));
                put(gi, nfi, f, q);
                centre(f, q);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j <= njc; ++j)
            for (int i = 0; i < nic; ++i) {
                const size_t f = (size_t)i + (size_t)nic * (j + (size_t)(njc + 1) * k);
                const geom::Quad q = geom::quad(V(i + 1, j, k), V(i, j, k), V(i, j, k + 1), V(i + 1, j, k + 1));
                put(gj, nfj, f, q);
                centre(nfi + f, q);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i < nic; ++i) {
                const size_t f = (size_t)i + (size_t)nic * (j + (size_t)njc * k);
                const geom::Quad q = geom::quad(V(i, j, k), V(i + 1, j, k), V(i + 1, j + 1, k), V(i, j + 1, k));
                put(gk, nfk, f, q);
                centre(nfi + nfj + f, q);
            }
}

// Primitive stage state from IO variables (rho, e, YB, vel from the flow file).
void prim_from_io(const real* io, int n, std::vector<real>& w)
{
    w.assign((size_t)6 * n, 0.0);
    for (int c = 0; c < n; ++c) {
        w[c] = io[6 * (size_t)n + c];              // rho
        w[(size_t)n + c] = io[7 * (size_t)n + c];  // e
        w[2 * (size_t)n + c] = io[8 * (size_t)n + c];  // YB
        w[3 * (size_t)n + c] = io[10 * (size_t)n + c];
        w[4 * (size_t)n + c] = io[11 * (size_t)n + c];
        w[5 * (size_t)n + c] = io[12 * (size_t)n + c];
    }
}

void conserved_from_prim(const std::vector<real>& w, int n, std::vector<real>& u)
{
    u.assign((size_t)6 * n, 0.0);
    for (int c = 0; c < n; ++c) {
        Prim p{w[c], w[(size_t)n + c], w[2 * (size_t)n + c], w[3 * (size_t)n + c],
               w[4 * (size_t)n + c], w[5 * (size_t)n + c]};
        real U[6];
        encode(p, U);  // simulate.cu:196-203 via flow.cu:227-243
        for (int q = 0; q < 6; ++q) u[(size_t)q * n + c] = U[q];
    }
}

void upload_state(chkn_sim* s, HostBlock& b, const real* io)
{
    std::vector<real> w, u;
    prim_from_io(io, b.n, w);
    size_t bytes = sizeof(real) * 6 * (size_t)b.n;
    int nbuf = b.active ? 3 : 1;
    for (int r = 0; r < nbuf; ++r) HIPCHK(hipMemcpy(b.W + (size_t)r * 6 * b.n, w.data(), bytes, hipMemcpyHostToDevice));
    if (b.active) {
        conserved_from_prim(w, b.n, u);
        HIPCHK(hipMemcpy(b.U, u.data(), bytes, hipMemcpyHostToDevice));
    }
}

// k-chunk lengths of a single block's plane-major tile table that give every XCD the same
// planes and a whole number of rounds of its resident work-groups: the planes in kXcds equal
// slabs (one per XCD's range of the table), each slab in p layers of near-equal length,
// longer first, for the smallest p (chunks of at most 8 planes) with cols * p a multiple of
// the XCD's slots.  Empty when the planes do not divide into kXcds slabs or no p fits.  A
// launch of, say, 5.33 rounds leaves the XCDs two-thirds idle through its last round.
static std::vector<int> xcd_layer_lengths(int nkc, long long cols, long long slots_per_xcd)
{
    std::vector<int> len;
    if (nkc % kXcds != 0 || slots_per_xcd <= 0) return len;
    const int pg = nkc / kXcds;
    for (int p = std::max(1, (pg + 7) / 8); p <= std::min(8, pg); ++p) {
        if ((cols * p) % slots_per_xcd != 0) continue;
        for (int g = 0; g < kXcds; ++g)
            for (int c = 0; c < p; ++c) len.push_back(pg / p + (c < pg % p ? 1 : 0));
        return len;
    }
    return len;
}

void build_tiles(chkn_sim* s)
{
    std::vector<Tile> tiles;
    const int TI = s->opts.tile_i, TJ = s->opts.tile_j;
    int nactive = 0;
    for (auto& o : s->blocks) nactive += o.active ? 1 : 0;
    for (size_t ib = 0; ib < s->blocks.size(); ++ib) {
        HostBlock& b = s->blocks[ib];
        b.tile_first = (int)tiles.size();
        if (b.active) {
            int nti = (b.nic + TI - 1) / TI, ntj = (b.njc + TJ - 1) / TJ;
            int kc = s->opts.tile_k > 0 ? s->opts.tile_k : s->kc;
            if (s->opts.tile_k <= 0) {
                // Default k-chunk: aim for about 2048 work-groups over the whole launch (one
                // plane per work-group for jobs too small to fill the device otherwise).
                long long cols = 0;
                for (auto& o : s->blocks)
                    if (o.active) cols += (long long)((o.nic + TI - 1) / TI) * ((o.njc + TJ - 1) / TJ) * o.nkc;
                long long want = std::max(1LL, cols / 2048);
                kc = (int)std::min<long long>(want, 32);
            }
            std::vector<int> layers;
            if (s->opts.tile_k <= 0 && s->opts.tile_order == 0 && nactive == 1) {
                const long long slots = (long long)(s->ncu / kXcds) * stage_slots_per_cu(TI, TJ, s->x_order);
                layers = xcd_layer_lengths(b.nkc, (long long)nti * ntj, slots);
            }
            if (!layers.empty()) {
                int k0 = 0;
                for (int nk : layers) {
                    for (int tj = 0; tj < ntj; ++tj)
                        for (int ti = 0; ti < nti; ++ti) tiles.push_back(Tile{(int)ib, ti * TI, tj * TJ, k0, nk});
                    k0 += nk;
                }
                if (k0 != b.nkc) throw std::logic_error("k-chunk layers do not cover the block");
            } else if (s->opts.tile_order == 1) {
                for (int tj = 0; tj < ntj; ++tj)
                    for (int ti = 0; ti < nti; ++ti)
                        for (int k0 = 0; k0 < b.nkc; k0 += kc)
                            tiles.push_back(Tile{(int)ib, ti * TI, tj * TJ, k0, std::min(kc, b.nkc - k0)});
            } else {
                for (int k0 = 0; k0 < b.nkc; k0 += kc)
                    for (int tj = 0; tj < ntj; ++tj)
                        for (int ti = 0; ti < nti; ++ti)
                            tiles.push_back(Tile{(int)ib, ti * TI, tj * TJ, k0, std::min(kc, b.nkc - k0)});
            }
        }
        b.tile_count = (int)tiles.size() - b.tile_first;
    }
    dfree(s->d_tiles);
    s->d_tiles = nullptr;
    s->ntiles = (int)tiles.size();
    if (!tiles.empty()) {
        HIPCHK(hipMalloc(&s->d_tiles, sizeof(Tile) * tiles.size()));
        HIPCHK(hipMemcpy(s->d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
    }
    for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
}

// Control-area views for segment slot q: dt[S][N], sigmin[S][N], run[S][N], bad[S][3],
// chem[S][N] with S = kMaxSeg step slots and N instances.
namespace {
size_t ctl_off_sig(int N) { return sizeof(real) * kMaxSeg * (size_t)N; }
size_t ctl_off_run(int N) { return ctl_off_sig(N) + sizeof(unsigned long long) * kMaxSeg * (size_t)N; }
size_t ctl_off_bad(int N) { return ctl_off_run(N) + sizeof(int) * kMaxSeg * (size_t)N; }
size_t ctl_off_chem(int N)
{
    const size_t o = ctl_off_bad(N) + sizeof(int) * 3 * kMaxSeg;
    return (o + 7) & ~size_t(7);
}
}  // namespace

real* ctl_dt(char* base, int q, int N) { return reinterpret_cast<real*>(base) + (size_t)q * N; }
unsigned long long* ctl_sig(char* base, int q, int N)
{
    return reinterpret_cast<unsigned long long*>(base + ctl_off_sig(N)) + (size_t)q * N;
}
int* ctl_run(char* base, int q, int N) { return reinterpret_cast<int*>(base + ctl_off_run(N)) + (size_t)q * N; }
int* ctl_bad(char* base, int q, int N) { return reinterpret_cast<int*>(base + ctl_off_bad(N)) + 3 * q; }
real* ctl_chem(char* base, int q, int N) { return reinterpret_cast<real*>(base + ctl_off_chem(N)) + (size_t)q * N; }
size_t ctl_size(int N) { return ctl_off_chem(N) + sizeof(real) * kMaxSeg * (size_t)N; }

// The minimum signal time from its bit pattern (non-negative reals order as their patterns).
real sig_value(unsigned long long bits)
{
    if (bits == ~0ull) return std::numeric_limits<real>::max();
    real m;
    if (sizeof(real) == sizeof(unsigned long long)) {
        std::memcpy(&m, &bits, sizeof m);
    } else {
        const unsigned int lo = (unsigned int)bits;
        std::memcpy(&m, &lo, sizeof m);
    }
    return m;
}

// exp(-alpha dt) of the A->B reaction (gas.cu:108), the same for every cell of an instance.
void fill_chem(chkn_sim* s, int q)
{
    const int N = (int)s->inst.size();
    real* c = ctl_chem(s->h_ctl, q, N);
    const real* dt = ctl_dt(s->h_ctl, q, N);
    for (int qi = 0; qi < N; ++qi) c[qi] = s->inst[qi].reacting ? std::exp(-kAlpha * dt[qi]) : real(-1);
}

// Size limits of one block.  Kernels address one component of a per-block array (a W slot,
// U, L, a direction's face geometry, the vertices) with an unsigned 32-bit byte offset from
// its base, and form cell, face and vertex indices in 32-bit ints; component bases use 64-bit
// arithmetic.  So every per-direction face count and the vertex count must stay below
// 2^32 / sizeof(real) (2^29 in FP64), and for viscous jobs the 4-value face records of all
// three directions as well.  Null when the block fits.
const char* block_dims_error(long long nic, long long njc, long long nkc, bool viscous)
{
    const long long lim = (1LL << 32) / (long long)sizeof(real);
    const long long nfi = (nic + 1) * njc * nkc, nfj = nic * (njc + 1) * nkc, nfk = nic * njc * (nkc + 1);
    const long long nv = (nic + 1) * (njc + 1) * (nkc + 1);
    if (std::max({nfi, nfj, nfk}) >= lim || nv >= lim || (viscous && 4 * (nfi + nfj + nfk) >= lim))
        return "block too large for 32-bit array offsets";
    return nullptr;
}

void check_desc(const chkn_desc* d)
{
    if (d->nblocks <= 0) throw std::invalid_argument("no blocks");
    if (d->gas_model && !gas_model_matches(d->gas_model)) throw std::invalid_argument("Incorrect gas model data");
    if (d->flux_calc != 0 && d->flux_calc != 1) throw std::invalid_argument("flux_calc must be 0 (ausmdv) or 1 (sbp_asf)");
    if (d->x_order != 1 && d->x_order != 2) throw std::invalid_argument("x_order must be 1 or 2");
    if (d->n_cfl <= 0) throw std::invalid_argument("need at least one cfl value");
    if (d->cfl_count <= 0) throw std::invalid_argument("cfl_count must be positive");
    if (d->time_scheme != CHKN_TIME_RK3 && d->time_scheme != CHKN_TIME_PC)
        throw std::invalid_argument("time_scheme must be 0 (rk3) or 1 (predictor_corrector)");
}

// Set up instances from job descriptions: validation, inflow states, geometry pre-pass,
// device buffers (initialize_simulation, simulate.cu:55-155, for each instance).
void setup(chkn_sim* s, const chkn_desc* descs, int ninst)
{
    if (ninst <= 0) throw std::invalid_argument("no instances");
    for (int q = 0; q < ninst; ++q) check_desc(&descs[q]);
    s->x_order = descs[0].x_order;
    s->flux_calc = descs[0].flux_calc;
    s->viscous = descs[0].viscous != 0;
    s->time_scheme = descs[0].time_scheme;
    for (int q = 1; q < ninst; ++q) {
        if (descs[q].time_scheme != s->time_scheme) throw std::invalid_argument("instances must share time_scheme");
        if (descs[q].x_order != s->x_order) throw std::invalid_argument("instances must share x_order");
        if (descs[q].flux_calc != s->flux_calc) throw std::invalid_argument("instances must share flux_calc");
        if ((descs[q].viscous != 0) != s->viscous) throw std::invalid_argument("instances must share viscous");
        if (descs[q].cfl_count != descs[0].cfl_count) throw std::invalid_argument("instances must share cfl_count");
    }
    s->inst.resize(ninst);
    int nblocks_total = 0;
    for (int q = 0; q < ninst; ++q) nblocks_total += descs[q].nblocks;
    s->blocks.resize(nblocks_total);
    int fs_base = 0, blk_base = 0;
    std::vector<real> inflow_pTYB;  // configured p, T, YB of every flow state
    for (int qi = 0; qi < ninst; ++qi) {
        const chkn_desc* d = &descs[qi];
        Instance& I = s->inst[qi];
        I.first_block = blk_base;
        I.nblocks = d->nblocks;
        I.nib = d->nib; I.njb = d->njb; I.nkb = d->nkb;
        I.blk_ids.assign(d->blk_ids, d->blk_ids + (size_t)d->nib * d->njb * d->nkb);
        I.cfl_count = d->cfl_count; I.max_step = d->max_step; I.print_count = d->print_count;
        I.reacting = d->reacting != 0;
        s->any_reacting = s->any_reacting || I.reacting;
        // Configuration values are held as real, as the reference holds them as number
        // (config.cu:494-504).
        I.max_time = (real)d->max_time; I.dt_init = (real)d->dt_init;
        I.cfl_t = to_real(d->cfl_times, d->n_cfl);
        I.cfl_v = to_real(d->cfl_values, d->n_cfl);
        if (d->n_dt_plot > 0) {
            I.plot_t = to_real(d->dt_plot_times, d->n_dt_plot);
            I.plot_v = to_real(d->dt_plot_values, d->n_dt_plot);
        } else {
            I.plot_t.assign(1, real(0));
            I.plot_v.assign(1, I.max_time);
        }
        // Inflow states: update_from_pT of the configured p, T, YB (config.cu:398-406).
        for (int q = 0; q < d->n_flow_states; ++q) {
            const std::vector<real> f = to_real(d->flow_states + 6 * q, 6);
            real p = f[0], T = f[1], YB = f[2];
            real e = T * kCv - YB * kQ;   // gas.cu:123
            real rho = p / (T * kR);      // gas.cu:124
            real w[6] = {rho, e, YB, f[3], f[4], f[5]};
            s->inflow_w.insert(s->inflow_w.end(), w, w + 6);
            inflow_pTYB.insert(inflow_pTYB.end(), f.begin(), f.begin() + 3);
        }
        size_t ooff = 0, twall_used = 0;
        for (int lb = 0; lb < d->nblocks; ++lb) {
            HostBlock& b = s->blocks[blk_base + lb];
            b.inst = qi;
            b.bi = d->blk_ijk[3 * lb]; b.bj = d->blk_ijk[3 * lb + 1]; b.bk = d->blk_ijk[3 * lb + 2];
            b.nic = d->blk_dims[3 * lb]; b.njc = d->blk_dims[3 * lb + 1]; b.nkc = d->blk_dims[3 * lb + 2];
            b.active = d->active[lb] != 0;
            for (int f = 0; f < 6; ++f) {
                b.bc[f] = d->bc_codes[6 * lb + f];
                b.bc_fs[f] = d->bc_fs[6 * lb + f] + fs_base;
                b.twall[f] = d->bc_twall ? (real)d->bc_twall[6 * lb + f] : real(0);
                b.bc_fun[f] = d->bc_fun ? d->bc_fun[6 * lb + f] : 0;
                b.twall_faces[f].clear();
                if (b.bc[f] == CHKN_BC_WALL_NO_SLIP_FIXED_T && std::isnan(b.twall[f])) {
                    // TWall per face, in block and side order (block.cu:566-611).
                    const size_t nf = f < 2 ? (size_t)b.njc * b.nkc : (f < 4 ? (size_t)b.nic * b.nkc : (size_t)b.nic * b.njc);
                    if (!d->twall_faces || twall_used + nf > (size_t)d->n_twall_faces)
                        throw std::invalid_argument("TWall given per face, but twall_faces holds too few values");
                    b.twall_faces[f] = to_real(d->twall_faces + twall_used, nf);
                    twall_used += nf;
                }
                if (b.active && b.bc[f] == CHKN_BC_INFLOW_FUNCTION) {  // bcs.cu:321-352
                    if (b.bc_fun[f] < 1 || b.bc_fun[f] > 3)
                        throw std::invalid_argument("inflow_function side without a known fun_name");
                    if (b.bc_fun[f] == 2 && !d->bl_profile)
                        throw std::invalid_argument("Boundary-layer file is not present in current directory.");
                }
            }
            // block.cu:282-284 and its j/k twins.
            if (b.nic < 2) throw std::invalid_argument("Too few cells in the i-index direction.");
            if (b.njc < 2) throw std::invalid_argument("Too few cells in the j-index direction.");
            if (b.nkc < 2) throw std::invalid_argument("Too few cells in the k-index direction.");
            if (const char* e = block_dims_error(b.nic, b.njc, b.nkc, s->viscous)) throw std::invalid_argument(e);
            const long long n = (long long)b.nic * b.njc * b.nkc;
            b.n = (int)n;
            if (b.bi < 0 || b.bi >= d->nib || b.bj < 0 || b.bj >= d->njb || b.bk < 0 || b.bk >= d->nkb)
                throw std::invalid_argument("block position outside the block array");
            if (I.blk_ids[((size_t)b.bi * d->njb + b.bj) * d->nkb + b.bk] != lb)
                throw std::invalid_argument("blk_id=" + std::to_string(lb) + " is inconsistent");
            for (int f = 0; f < 6; ++f) {
                if (b.active && b.bc[f] == CHKN_BC_INFLOW &&
                    (d->bc_fs[6 * lb + f] < 0 || d->bc_fs[6 * lb + f] >= d->n_flow_states))
                    throw std::invalid_argument("inflow flow_state_index out of range");
            }
            b.io0 = to_real(d->iovars + ooff, (size_t)13 * b.n);
            ooff += (size_t)13 * b.n;
        }
        for (auto& id : I.blk_ids) {
            if (id >= d->nblocks) throw std::invalid_argument("blk_ids entry beyond nblocks");
            if (id >= 0) id += blk_base;
        }
        fs_base += d->n_flow_states;
        blk_base += d->nblocks;
    }
    if (s->inflow_w.empty()) s->inflow_w.assign(6, real(0));
    // Exchange partners (bcs.cu:145-252, periodic wrap at the array edges).
    std::vector<std::array<int, 6>> partner(s->blocks.size());
    for (size_t ib = 0; ib < s->blocks.size(); ++ib) {
        HostBlock& b = s->blocks[ib];
        const Instance& I = s->inst[b.inst];
        for (int f = 0; f < 6; ++f) {
            partner[ib][f] = -1;
            if (!b.active || b.bc[f] != CHKN_BC_EXCHANGE) continue;
            int oi = b.bi + (f == 1) - (f == 0), oj = b.bj + (f == 3) - (f == 2), ok = b.bk + (f == 5) - (f == 4);
            oi = (oi + I.nib) % I.nib; oj = (oj + I.njb) % I.njb; ok = (ok + I.nkb) % I.nkb;
            int id = I.blk_ids[((size_t)oi * I.njb + oj) * I.nkb + ok];
            if (id < 0 || id >= I.first_block + I.nblocks)
                throw std::invalid_argument("exchange boundary without a neighbour block");
            HostBlock& o = s->blocks[id];
            bool okd = (f < 2) ? (o.njc == b.njc && o.nkc == b.nkc)
                       : (f < 4) ? (o.nic == b.nic && o.nkc == b.nkc) : (o.nic == b.nic && o.njc == b.njc);
            if (!okd) throw std::invalid_argument("exchange neighbour has mismatched face dimensions");
            partner[ib][f] = id;
        }
    }
    // Reverse partners: the block whose side 2d+1 (2d) exchanges into side 2d (2d+1).
    std::vector<std::array<int, 6>> rpart(s->blocks.size());
    for (auto& r : rpart) r.fill(-1);
    for (size_t ib = 0; ib < s->blocks.size(); ++ib)
        for (int f = 0; f < 6; ++f)
            if (partner[ib][f] >= 0 && s->blocks[partner[ib][f]].active)
                rpart[partner[ib][f]][f ^ 1] = (int)ib;
    // Device.
    HIPCHK(hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking));
    s->stream = s->own_stream;
    HIPCHK(hipEventCreate(&s->ev0));
    HIPCHK(hipEventCreate(&s->ev1));
    std::vector<BlockDev> bdev(s->blocks.size());
    for (int qi = 0; qi < ninst; ++qi) {
        const chkn_desc* d = &descs[qi];
        size_t voff = 0;
        bl::Profile profile;  // the instance's laminar boundary-layer profile (config.cu:37-60)
        if (d->bl_profile) profile.set(to_real(d->bl_profile, bl::kProfileValues).data());
        for (int lb = 0; lb < d->nblocks; ++lb) {
            int ib = s->inst[qi].first_block + lb;
            HostBlock& b = s->blocks[ib];
            std::vector<real> gi, gj, gk, volinv, len, fpos;
            bool want_fpos = s->viscous && b.active;
            for (int f = 0; f < 6; ++f) want_fpos = want_fpos || (b.active && b.bc[f] == CHKN_BC_INFLOW_FUNCTION);
            const size_t nv = (size_t)3 * (b.nic + 1) * (b.njc + 1) * (b.nkc + 1);
            const std::vector<real> vxyz = to_real(d->vertices + voff, nv);
            block_geometry(b, vxyz.data(), gi, gj, gk, volinv, len, want_fpos ? &fpos : nullptr);
            voff += nv;
            size_t n = (size_t)b.n, sz6 = sizeof(real) * 6 * n;
            dalloc(s, (void**)&b.W, sz6 * (b.active ? 3 : 1));
            BlockDev& bd = bdev[ib];
            std::memset(&bd, 0, sizeof bd);
            bd.nic = b.nic; bd.njc = b.njc; bd.nkc = b.nkc; bd.active = b.active ? 1 : 0;
            bd.inst = b.inst;
            bd.n = b.n;
            bd.nfi = (long long)(b.nic + 1) * b.njc * b.nkc;
            bd.nfj = (long long)b.nic * (b.njc + 1) * b.nkc;
            bd.nfk = (long long)b.nic * b.njc * (b.nkc + 1);
            bd.W = b.W;
            bd.wstride = b.active ? (long long)(6 * n) : 0;
            if (b.active) {
                dalloc(s, (void**)&b.U, sz6);
                dalloc(s, (void**)&b.L, sz6);
                HIPCHK(hipMemset(b.L, 0, sz6));
                dalloc(s, (void**)&b.gi, sizeof(real) * gi.size());
                dalloc(s, (void**)&b.gj, sizeof(real) * gj.size());
                dalloc(s, (void**)&b.gk, sizeof(real) * gk.size());
                dalloc(s, (void**)&b.volinv, sizeof(real) * n);
                dalloc(s, (void**)&b.len, sizeof(real) * 3 * n);
                HIPCHK(hipMemcpy(b.gi, gi.data(), sizeof(real) * gi.size(), hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(b.gj, gj.data(), sizeof(real) * gj.size(), hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(b.gk, gk.data(), sizeof(real) * gk.size(), hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(b.volinv, volinv.data(), sizeof(real) * n, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(b.len, len.data(), sizeof(real) * 3 * n, hipMemcpyHostToDevice));
                bd.U = b.U; bd.L = b.L; bd.gi = b.gi; bd.gj = b.gj; bd.gk = b.gk;
                bd.volinv = b.volinv; bd.len = b.len;
                // Vertices as [3][nv] (x, then y, then z) for the frames formed on the device.
                const size_t nvert = nv / 3;
                std::vector<real> soa(nv);
                for (size_t v = 0; v < nvert; ++v)
                    for (int c = 0; c < 3; ++c) soa[c * nvert + v] = vxyz[3 * v + c];
                dalloc(s, (void**)&b.vtx, sizeof(real) * nv);
                HIPCHK(hipMemcpy(b.vtx, soa.data(), sizeof(real) * nv, hipMemcpyHostToDevice));
                bd.vtx = b.vtx;
                bd.nv = (long long)nvert;
            }
            for (int f = 0; f < 6; ++f) {
                bd.rp[f] = rpart[ib][f];
                bd.bc[f] = b.bc[f];
                bd.bc_arg[f] = (b.bc[f] == CHKN_BC_EXCHANGE) ? partner[ib][f] : b.bc_fs[f];
            }
            upload_state(s, b, b.io0.data());
            split_block_setup(s, b, bd, fpos);
            split_block_gas(s, b, bd);
            side_data_setup(b, bd, fpos, d->bl_profile ? &profile : nullptr);
        }
    }
    s->bdev = bdev;
    dalloc(s, (void**)&s->d_blocks, sizeof(BlockDev) * bdev.size());
    upload_blocks(s);
    dalloc(s, (void**)&s->d_inflow, sizeof(real) * s->inflow_w.size());
    HIPCHK(hipMemcpy(s->d_inflow, s->inflow_w.data(), sizeof(real) * s->inflow_w.size(), hipMemcpyHostToDevice));
    split_inflow_gas(s, inflow_pTYB);
    split_tables(s);
    const int N = ninst;
    s->ctl_bytes = ctl_size(N);
    dalloc(s, (void**)&s->d_ctl, s->ctl_bytes);
    HIPCHK(hipHostMalloc(&s->h_ctl, s->ctl_bytes, hipHostMallocDefault));
    std::memset(s->h_ctl, 0, s->ctl_bytes);
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    s->ncu = ncu;
    s->grid_limit = ncu * 2;
    build_tiles(s);
    // SimState at start (simulate.cu:132-150, :190-191).
    for (int qi = 0; qi < ninst; ++qi) {
        Instance& I = s->inst[qi];
        I.t = (real)descs[qi].t0;
        I.dt = I.dt_init;
        I.step = 0;
        I.next_plot_indx = descs[qi].start_tindx + 1;
        I.t_plot = I.t + schedule_get(I.plot_t, I.plot_v, I.t);
        I.steps_since_last_plot = 0;
        I.have_sig = false;
    }
    s->gstep = 0;
    s->fresh = true;
}

int stages_per_step(const chkn_sim* s) { return s->time_scheme == CHKN_TIME_PC ? 2 : 3; }

StageArgs stage_args(chkn_sim* s, int q)
{
    const int N = (int)s->inst.size();
    StageArgs a;
    a.blocks = s->d_blocks;
    a.tiles = s->d_tiles;
    a.inflow = s->d_inflow;
    a.ctl.dt = ctl_dt(s->d_ctl, q, N);
    a.ctl.sigmin_bits = ctl_sig(s->d_ctl, q, N);
    a.ctl.run = ctl_run(s->d_ctl, q, N);
    a.ctl.bad = ctl_bad(s->d_ctl, q, N);
    a.ctl.chem = ctl_chem(s->d_ctl, q, N);
    a.tile_base = 0;
    a.ntiles = 0;
    a.win = a.wout = 0;
    return a;
}

// Launch the 3*K stages of a segment; fresh: the segment starts from the state read from
// IO variables, whose first stage reads the file gas states.
void launch_segment(chkn_sim* s, int K, bool dtred_last, bool fresh)
{
    for (int q = 0; q < K; ++q) {
        StageArgs a = stage_args(s, q);
        // Stages 1, 2, 3 (TVD-RK3) or 1, 4 (predictor-corrector) in a ring of W slots:
        // each reads the last stage's output, the final stage writes slot 0.  On a tape the
        // step's states follow the previous step's.
        const int nst = stages_per_step(s);
        if (s->tape_tab) a.blocks = s->tape_tab;
        for (int pos = 1; pos <= nst; ++pos) {
            const int stage = (s->time_scheme == CHKN_TIME_PC && pos == 2) ? 4 : pos;
            const bool last = pos == nst;
            a.win = pos - 1;
            a.wout = pos % nst;
            if (s->tape_tab) {
                a.win = s->tape_base + nst * q + pos - 1;
                a.wout = s->tape_base + nst * q + pos;
            }
            bool dtred = last && dtred_last && (q == K - 1);
            // With reacting instances the reduction follows the reaction.
            launch_stage_form(s, a, stage, dtred && !s->any_reacting, fresh && q == 0 && pos == 1,
                              s->stream);
            if (last && s->any_reacting)
                HIPCHK(launch_chemistry(a.blocks, s->d_split_cells, s->nsplit_cells, a.ctl, a.wout,
                                        dtred, s->stream));
        }
    }
}

long long launches_per_segment(chkn_sim* s, int K, bool fresh)
{
    const long long per_stage = launches_per_stage(s, false);
    long long n = (long long)stages_per_step(s) * K * per_stage + (s->any_reacting ? K : 0);
    if (fresh) n += launches_per_stage(s, true) - per_stage;  // first stage in face-list form
    return n;
}

// run[q][inst] and dt[q][inst] must already be in h_ctl.
void run_segment(chkn_sim* s, int K, bool dtred_last)
{
    const int N = (int)s->inst.size();
    for (int q = 0; q < K; ++q) {
        unsigned long long* sg = ctl_sig(s->h_ctl, q, N);
        for (int qi = 0; qi < N; ++qi) sg[qi] = ~0ull;
        int* bd = ctl_bad(s->h_ctl, q, N);
        bd[0] = bd[1] = bd[2] = 0;
    }
    HIPCHK(hipMemcpyAsync(s->d_ctl, s->h_ctl, s->ctl_bytes, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipEventRecord(s->ev0, s->stream));
    const bool fresh = fresh_needs_file_gas(s);
    // The null stream cannot be captured; tape segments are launched directly.
    if (s->opts.use_graph && s->stream != nullptr && !s->tape_tab) {
        GraphKey key{K, dtred_last, fresh};
        auto it = s->graphs.find(key);
        if (it == s->graphs.end()) {
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            launch_segment(s, K, dtred_last, fresh);
            HIPCHK(hipStreamEndCapture(s->stream, &g));
            hipGraphExec_t ge;
            HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            it = s->graphs.emplace(key, ge).first;
        }
        HIPCHK(hipGraphLaunch(it->second, s->stream));
    } else {
        launch_segment(s, K, dtred_last, fresh);
    }
    s->last_launches += launches_per_segment(s, K, fresh);
    HIPCHK(hipEventRecord(s->ev1, s->stream));
    HIPCHK(hipMemcpyAsync(s->h_ctl, s->d_ctl, s->ctl_bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms += (double)ms;
}

// Minimum signal time of each instance's current state when no stage-3 launch reduced it
// (first CFL step after the state was replaced).
void signal_reduction(chkn_sim* s)
{
    const int N = (int)s->inst.size();
    unsigned long long* sg = ctl_sig(s->h_ctl, 0, N);
    for (int qi = 0; qi < N; ++qi) sg[qi] = ~0ull;
    HIPCHK(hipMemcpyAsync(ctl_sig(s->d_ctl, 0, N), sg, sizeof(unsigned long long) * N,
                          hipMemcpyHostToDevice, s->stream));
    for (size_t ib = 0; ib < s->blocks.size(); ++ib) {
        HostBlock& b = s->blocks[ib];
        if (!b.active) continue;
        HIPCHK(launch_signal_min(s->d_blocks + ib, b.n, ctl_sig(s->d_ctl, 0, N) + b.inst, s->stream));
    }
    HIPCHK(hipMemcpyAsync(sg, ctl_sig(s->d_ctl, 0, N), sizeof(unsigned long long) * N,
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int qi = 0; qi < N; ++qi) {
        s->inst[qi].sigmin = sig_value(sg[qi]);
        s->inst[qi].have_sig = true;
    }
}

int write_instance_flow(chkn_sim* s, int qi, int tindx, bool binary);

void fill_report(chkn_sim* s, int qi, chkn_report* rep, int done)
{
    if (!rep) return;
    const Instance& I = s->inst[qi];
    rep->step = I.step; rep->t = (double)I.t; rep->dt = (double)I.dt;
    rep->bad_stage = 0; rep->bad_count = 0; rep->steps_done = done;
}

// The march's progress line (simulate.cu:311-326): step, time, dt, CFL, wall clock since the
// march started (whole milliseconds, as the reference counts them) and its estimates of the
// wall clock to max_time and to max_step.
void print_progress(const chkn_sim* s, const Instance& I, int qi)
{
    const long long ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                             std::chrono::steady_clock::now() - s->march_start).count();
    const double wc = (double)ms / 1000.0;
    const double per_step = wc / I.step;
    const double to_ft = (double)(I.max_time - I.t) / (double)I.dt * per_step;
    const double to_ms = (I.max_step - I.step) * per_step;
    if (s->inst.size() > 1) std::printf("[instance %d] ", qi);
    std::printf("Step=%d t=%.3e dt=%.3e cfl=%.3f WC=%.3fs WCtFT=%.3fs WCtMS=%.3fs\n", I.step, (double)I.t,
                (double)I.dt, (double)schedule_get(I.cfl_t, I.cfl_v, I.t), wc, to_ft, to_ms);
    std::fflush(stdout);
}

// Advance up to nsteps lockstep steps; when plot_writes (the march), stop segments at snapshot
// times and at progress-line steps, write snapshots and print the progress lines.  Restates
// the loop of simulate.cu:205-335 for every instance.
int do_steps(chkn_sim* s, int nsteps, bool plot_writes, bool binary, chkn_report* rep)
{
    if (s->failed) return set_error(s, CHKN_ERR_STATE, "simulation is in a failed state");
    s->last_ms = 0.0;
    s->last_launches = 0;
    const int N = (int)s->inst.size();
    const int cfl_count = s->inst[0].cfl_count;
    int done = 0;
    auto any_running = [&] {
        for (auto& I : s->inst) if (I.running()) return true;
        return false;
    };
    while (done < nsteps && any_running()) {
        if (s->gstep > 0 && (s->gstep % cfl_count) == 0) {  // simulate.cu:208-221
            bool need = false;
            for (auto& I : s->inst) if (I.running() && !I.have_sig) need = true;
            if (need) signal_reduction(s);
            for (auto& I : s->inst) {
                if (!I.running()) continue;
                real cfl = schedule_get(I.cfl_t, I.cfl_v, I.t);
                real smallest = cfl * I.sigmin;
                I.dt = std::fmin(real(1.5) * I.dt, smallest);
                I.have_sig = false;
            }
        }
        // Segment: constant dt until the next CFL step, bounded by the run limits.
        int to_cfl = cfl_count - (s->gstep % cfl_count);
        int K = std::min({to_cfl, nsteps - done, kMaxSeg});
        std::vector<int> nrun(N, 0);
        int Kuse = 0;
        for (int qi = 0; qi < N; ++qi) {
            Instance& I = s->inst[qi];
            real tt = I.t;
            int st = I.step, q = 0;
            while (q < K && st < I.max_step && tt < I.max_time) {
                tt += I.dt;
                ++st;
                ++q;
                if (plot_writes && tt >= I.t_plot) break;
            }
            nrun[qi] = q;
            Kuse = std::max(Kuse, q);
        }
        if (plot_writes)  // a snapshot or a progress line of any instance ends the segment for all
            for (int qi = 0; qi < N; ++qi) {
                Instance& I = s->inst[qi];
                real tt = I.t;
                for (int q = 0; q < nrun[qi]; ++q) {
                    tt += I.dt;
                    const bool print = I.print_count > 0 && (I.step + q + 1) % I.print_count == 0;
                    if (tt >= I.t_plot || print) { Kuse = std::min(Kuse, q + 1); break; }
                }
            }
        K = Kuse;
        for (int q = 0; q < K; ++q) {
            real* dtv = ctl_dt(s->h_ctl, q, N);
            int* run = ctl_run(s->h_ctl, q, N);
            for (int qi = 0; qi < N; ++qi) {
                dtv[qi] = s->inst[qi].dt;
                run[qi] = (q < nrun[qi]) ? 1 : 0;
            }
            fill_chem(s, q);
        }
        bool dtred = ((s->gstep + K) % cfl_count) == 0;
        run_segment(s, K, dtred);
        for (int q = 0; q < K; ++q) {
            const int* bad = ctl_bad(s->h_ctl, q, N);
            for (int st = 0; st < 3; ++st)
                if (bad[st] > 0) {
                    // Steps before the failing one completed (the rest of the segment's
                    // launches ran on the failed state; the handle is left failed).
                    s->failed = true;
                    for (int qi = 0; qi < N; ++qi) {
                        Instance& I = s->inst[qi];
                        for (int r = 0; r < std::min(q, nrun[qi]); ++r) { I.t += I.dt; I.step += 1; }
                    }
                    s->gstep += q;
                    if (rep) { rep->bad_stage = st + 1; rep->bad_count = bad[st]; }
                    char msg[128];
                    std::snprintf(msg, sizeof msg, "Stage %d bad cell count: %d", st + 1, bad[st]);
                    set_error(s, bad[st], msg);
                    return bad[st];
                }
        }
        if (dtred) {
            const unsigned long long* sg = ctl_sig(s->h_ctl, K - 1, N);
            for (int qi = 0; qi < N; ++qi) {
                if (nrun[qi] < K) continue;
                s->inst[qi].sigmin = sig_value(sg[qi]);
                s->inst[qi].have_sig = true;
            }
        }
        if (s->dt_log)
            for (int q = 0; q < K; ++q)
                for (int qi = 0; qi < N; ++qi) s->dt_log->push_back(q < nrun[qi] ? (double)s->inst[qi].dt : 0.0);
        for (int qi = 0; qi < N; ++qi) {  // simulate.cu:306-308; the device ran min(nrun, K)
            Instance& I = s->inst[qi];
            nrun[qi] = std::min(nrun[qi], K);
            for (int q = 0; q < nrun[qi]; ++q) {
                I.t += I.dt;
                I.step += 1;
                I.steps_since_last_plot += 1;
            }
        }
        s->gstep += K;
        s->fresh = false;
        done += K;
        if (s->tape_tab) s->tape_base += stages_per_step(s) * K;
        if (plot_writes)
            for (int qi = 0; qi < N; ++qi) {  // simulate.cu:311-334
                Instance& I = s->inst[qi];
                if (nrun[qi] > 0 && I.print_count > 0 && I.step % I.print_count == 0) print_progress(s, I, qi);
                if (nrun[qi] > 0 && I.t >= I.t_plot) {
                    int rc = write_instance_flow(s, qi, I.next_plot_indx, binary);
                    if (rc != CHKN_OK) return rc;
                    I.steps_since_last_plot = 0;
                    I.next_plot_indx += 1;
                    I.t_plot = I.t + schedule_get(I.plot_t, I.plot_v, I.t);
                }
            }
    }
    fill_report(s, 0, rep, done);
    return CHKN_OK;
}

// The 13 IO variables of block ib in the build's real type.
void block_iovars(chkn_sim* s, int ib, real* out)
{
    HostBlock& b = s->blocks[ib];
    size_t n = (size_t)b.n;
    std::memcpy(out, b.pos.data(), sizeof(real) * 3 * n);
    std::memcpy(out + 3 * n, b.vol.data(), sizeof(real) * n);
    if (!b.active || s->fresh) {
        std::memcpy(out + 4 * n, b.io0.data() + 4 * n, sizeof(real) * 9 * n);
        return;
    }
    std::vector<real> w(6 * n);
    HIPCHK(hipMemcpy(w.data(), b.W, sizeof(real) * 6 * n, hipMemcpyDeviceToHost));
    for (size_t c = 0; c < n; ++c) {
        real rho = w[c], e = w[n + c], YB = w[2 * n + c];
        real T = eos_T(e, YB), p = eos_p(rho, T), a = eos_a(T);  // gas.cu:128-135
        out[4 * n + c] = p; out[5 * n + c] = T; out[6 * n + c] = rho; out[7 * n + c] = e;
        out[8 * n + c] = YB; out[9 * n + c] = a;
        out[10 * n + c] = w[3 * n + c]; out[11 * n + c] = w[4 * n + c]; out[12 * n + c] = w[5 * n + c];
    }
}

// write_flow_data (simulate.cu:158-179) for one instance.
int write_instance_flow(chkn_sim* s, int qi, int tindx, bool binary)
{
    Instance& I = s->inst[qi];
    if (I.job.empty()) return set_error(s, CHKN_ERR_STATE, "no job directory to write into");
    char buf[64];
    std::snprintf(buf, sizeof buf, "/flow/t%04d", tindx);
    std::filesystem::create_directories(I.job + buf);
    for (int lb = 0; lb < I.nblocks; ++lb) {
        HostBlock& b = s->blocks[I.first_block + lb];
        std::vector<real> io((size_t)13 * b.n);
        block_iovars(s, I.first_block + lb, io.data());
        write_flow(block_file(I.job, "flow", tindx, b.bi, b.bj, b.bk, binary), binary, io.data(), b.n);
    }
    append_times(I.job, tindx, I.t);
    return CHKN_OK;
}

int create_from(const chkn_desc* descs, int n, chkn_sim** out)
{
    if (!descs || !out || n <= 0) return set_error(nullptr, CHKN_ERR_ARG, "null argument");
    *out = nullptr;
    auto s = std::make_unique<chkn_sim>();
    try {
        setup(s.get(), descs, n);
    } catch (const HipError& e) {
        free_device(s.get());
        return set_error(nullptr, CHKN_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        free_device(s.get());
        return set_error(nullptr, CHKN_ERR_ARG, e.what());
    } catch (const std::exception& e) {
        free_device(s.get());
        return set_error(nullptr, CHKN_ERR_CONFIG, e.what());
    }
    *out = s.release();
    return CHKN_OK;
}

}  // namespace rt
}  // namespace chkn

using namespace chkn::rt;

extern "C" {

int chkn_create(const chkn_desc* desc, chkn_sim** out) { return create_from(desc, 1, out); }

int chkn_create_batch(const chkn_desc* descs, int ninst, chkn_sim** out)
{
    return create_from(descs, ninst, out);
}

// Console echo of the job configuration at verbosity > 0 (after config.cu:367-523).
static void echo_config(const JobConfig& c)
{
    auto list = [](const char* name, const std::vector<long long>& v) {
        std::printf("  %s: [", name);
        for (long long x : v) std::printf("%lld,", x);
        std::printf("]\n");
    };
    std::printf("  title: %s\n", c.title.c_str());
    std::printf("  nib: %d\n  njb: %d\n  nkb: %d\n", c.nib, c.njb, c.nkb);
    std::vector<long long> nics(c.nib), njcs(c.njb), nkcs(c.nkb);
    for (auto& b : c.blocks) { nics[b.i] = b.nic; njcs[b.j] = b.njc; nkcs[b.k] = b.nkc; }
    list("nics", nics); list("njcs", njcs); list("nkcs", nkcs);
    std::printf("  x_order=%d\n  t_order=%d\n  time_scheme=%s\n  flux_calc=%s\n", c.x_order, c.t_order,
                c.time_scheme ? "predictor_corrector" : "rk3", c.flux_calc ? "sbp_asf" : "ausmdv");
    std::printf("  dt_init=%g\n  max_time=%g\n  max_step=%d\n  cfl_count=%d\n  print_count=%d\n", c.dt_init,
                c.max_time, c.max_step, c.cfl_count, c.print_count);
    std::printf("  viscous=%d\n  reacting=%d\n", c.viscous ? 1 : 0, c.reacting ? 1 : 0);
    std::printf("Finished reading JSON file.\n");
}

int chkn_check_block_dims(int nic, int njc, int nkc, int viscous)
{
    if (nic < 2 || njc < 2 || nkc < 2) return set_error(nullptr, CHKN_ERR_ARG, "Too few cells in a block direction.");
    if (const char* e = block_dims_error(nic, njc, nkc, viscous != 0)) return set_error(nullptr, CHKN_ERR_ARG, e);
    return CHKN_OK;
}

int chkn_set_verbosity(int level)
{
    g_verbosity = level;
    return CHKN_OK;
}

long long chkn_device_bytes(chkn_sim* sim) { return sim ? sim->device_bytes : -1; }

int chkn_load_job(const char* job_dir, int tindx, int binary, chkn_sim** out)
{
    if (!job_dir || !out) return set_error(nullptr, CHKN_ERR_ARG, "null argument");
    *out = nullptr;
    std::string job(job_dir);
    JobConfig c;
    std::vector<double> verts, io, twall_faces;
    double t0 = 0.0;
    int next_plot = 1;
    try {
        if (!std::filesystem::is_directory(job))
            return set_error(nullptr, CHKN_ERR_IO, "Job directory is not present in current directory.");
        if (g_verbosity > 0) std::printf("Reading JSON config file.\n");
        c = read_config_json(job + "/config.json");
        if (g_verbosity > 0) echo_config(c);
        for (auto& b : c.blocks) {
            std::vector<double> v = read_grid(block_file(job, "grid", 0, b.i, b.j, b.k, binary != 0), binary != 0, b.nic, b.njc, b.nkc);
            verts.insert(verts.end(), v.begin(), v.end());
            std::vector<double> f = read_flow(block_file(job, "flow", tindx, b.i, b.j, b.k, binary != 0), binary != 0, b.nic * b.njc * b.nkc);
            io.insert(io.end(), f.begin(), f.end());
        }
        read_times(job, tindx, t0, next_plot);
        for (auto& b : c.blocks)
            for (int f = 0; f < 6; ++f)
                if (b.bc[f] == CHKN_BC_WALL_NO_SLIP_FIXED_T && std::isnan(b.bc_twall[f])) {
                    const int nf = f < 2 ? b.njc * b.nkc : (f < 4 ? b.nic * b.nkc : b.nic * b.njc);
                    std::vector<double> tw = read_twall(job, b, f, nf);
                    twall_faces.insert(twall_faces.end(), tw.begin(), tw.end());
                }
    } catch (const std::runtime_error& e) {
        std::string m = e.what();
        return set_error(nullptr, m == "Incorrect gas model data" ? CHKN_ERR_CONFIG : CHKN_ERR_IO, m);
    }
    std::vector<int> ijk, dims, active, bcc, bcf, bcfun;
    std::vector<double> twall;
    for (auto& b : c.blocks) {
        ijk.insert(ijk.end(), {b.i, b.j, b.k});
        dims.insert(dims.end(), {b.nic, b.njc, b.nkc});
        active.push_back(b.active ? 1 : 0);
        bcc.insert(bcc.end(), b.bc, b.bc + 6);
        bcf.insert(bcf.end(), b.bc_fs, b.bc_fs + 6);
        twall.insert(twall.end(), b.bc_twall, b.bc_twall + 6);
        bcfun.insert(bcfun.end(), b.bc_fun, b.bc_fun + 6);
    }
    chkn_desc d;
    std::memset(&d, 0, sizeof d);
    d.nblocks = (int)c.blocks.size(); d.nib = c.nib; d.njb = c.njb; d.nkb = c.nkb;
    d.blk_ids = c.blk_ids.data(); d.blk_ijk = ijk.data(); d.blk_dims = dims.data();
    d.active = active.data(); d.bc_codes = bcc.data(); d.bc_fs = bcf.data();
    d.n_flow_states = (int)c.flow_states.size() / 6; d.flow_states = c.flow_states.data();
    d.vertices = verts.data(); d.iovars = io.data();
    d.x_order = c.x_order; d.flux_calc = c.flux_calc; d.t_order = c.t_order;
    d.cfl_count = c.cfl_count; d.max_step = c.max_step; d.print_count = c.print_count;
    d.max_time = c.max_time; d.dt_init = c.dt_init;
    d.n_cfl = (int)c.cfl_times.size(); d.cfl_times = c.cfl_times.data(); d.cfl_values = c.cfl_values.data();
    d.n_dt_plot = (int)c.dt_plot_times.size(); d.dt_plot_times = c.dt_plot_times.data();
    d.dt_plot_values = c.dt_plot_values.data();
    d.t0 = t0; d.start_tindx = next_plot - 1;
    d.reacting = c.reacting; d.viscous = c.viscous;
    d.time_scheme = c.time_scheme;
    d.bc_twall = twall.data();
    d.bc_fun = bcfun.data();
    d.gas_model = c.gas_model.data();
    d.bl_profile = c.bl_profile.empty() ? nullptr : c.bl_profile.data();
    d.n_twall_faces = (int)twall_faces.size();
    d.twall_faces = twall_faces.empty() ? nullptr : twall_faces.data();
    int rc = chkn_create(&d, out);
    if (rc == CHKN_OK) (*out)->inst[0].job = job;
    return rc;
}

int chkn_step(chkn_sim* sim, int nsteps, chkn_report* rep)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    return guarded(sim, [&] { return do_steps(sim, nsteps, false, true, rep); });
}

int chkn_march(chkn_sim* sim, int binary, chkn_report* rep)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    for (auto& I : sim->inst)
        if (I.job.empty()) return set_error(sim, CHKN_ERR_STATE, "chkn_march needs a job directory");
    sim->march_start = std::chrono::steady_clock::now();  // simulate.cu:189
    if (g_verbosity > 0) std::printf("march_in_time start\n");
    const int rc = guarded(sim, [&] { return do_steps(sim, 0x7fffffff, true, binary != 0, rep); });
    if (g_verbosity > 0) std::printf("march_in_time end\n");
    return rc;
}

int chkn_finalize(chkn_sim* sim, int binary)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    if (sim->failed) return CHKN_OK;  // the reference skips finalize after a thrown error
    return guarded(sim, [&] {
        for (int qi = 0; qi < (int)sim->inst.size(); ++qi) {
            Instance& I = sim->inst[qi];
            if (I.steps_since_last_plot > 0 && !I.job.empty()) {
                int rc = write_instance_flow(sim, qi, I.next_plot_indx, binary != 0);
                if (rc != CHKN_OK) return rc;
            }
        }
        return CHKN_OK;
    });
}

int chkn_write_flow(chkn_sim* sim, int tindx, int binary)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    return guarded(sim, [&] {
        for (int qi = 0; qi < (int)sim->inst.size(); ++qi) {
            int rc = write_instance_flow(sim, qi, tindx, binary != 0);
            if (rc != CHKN_OK) return rc;
        }
        return CHKN_OK;
    });
}

int chkn_get_iovars(chkn_sim* sim, int blk, double* out)
{
    if (!sim || !out || blk < 0 || blk >= (int)sim->blocks.size()) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        std::vector<real> io((size_t)13 * sim->blocks[blk].n);
        block_iovars(sim, blk, io.data());
        std::copy(io.begin(), io.end(), out);
        return CHKN_OK;
    });
}

int chkn_get_conserved(chkn_sim* sim, int blk, double* out)
{
    if (!sim || !out || blk < 0 || blk >= (int)sim->blocks.size()) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        HostBlock& b = sim->blocks[blk];
        if (!b.active) { std::memset(out, 0, sizeof(double) * 6 * (size_t)b.n); return CHKN_OK; }
        std::vector<real> u((size_t)6 * b.n);
        HIPCHK(hipMemcpy(u.data(), b.U, sizeof(real) * u.size(), hipMemcpyDeviceToHost));
        std::copy(u.begin(), u.end(), out);
        return CHKN_OK;
    });
}

int chkn_set_iovars(chkn_sim* sim, int blk, const double* iovars)
{
    if (!sim || !iovars || blk < 0 || blk >= (int)sim->blocks.size()) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        HostBlock& b = sim->blocks[blk];
        // The handle is marked fresh below: every block's first stage then reads its held
        // IO variables' gas states.  Blocks that have stepped since they were set take their
        // current state as held state (gas states from the equation of state of W, which the
        // stage would compute anyway).
        if (!sim->fresh)
            for (size_t ib = 0; ib < sim->blocks.size(); ++ib) {
                HostBlock& o = sim->blocks[ib];
                if ((int)ib == blk || !o.active) continue;
                block_iovars(sim, (int)ib, o.io0.data());
                split_block_gas(sim, o, sim->bdev[ib]);
            }
        const std::vector<real> v = to_real(iovars + 4 * (size_t)b.n, 9 * (size_t)b.n);
        std::copy(v.begin(), v.end(), b.io0.begin() + 4 * (size_t)b.n);
        upload_state(sim, b, b.io0.data());
        split_block_gas(sim, b, sim->bdev[blk]);
        upload_blocks(sim);
        for (auto& kv : sim->graphs) (void)hipGraphExecDestroy(kv.second);
        sim->graphs.clear();
        sim->fresh = true;
        sim->failed = false;
        for (auto& I : sim->inst) I.have_sig = false;
        return CHKN_OK;
    });
}

int chkn_set_stream(chkn_sim* sim, void* stream)
{
    if (!sim || !sim->own_stream) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        HIPCHK(hipStreamSynchronize(sim->stream));  // work queued so far completes first
        sim->stream = static_cast<hipStream_t>(stream);
        return CHKN_OK;
    });
}

int chkn_reset_stream(chkn_sim* sim)
{
    if (!sim || !sim->own_stream) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        HIPCHK(hipStreamSynchronize(sim->stream));
        sim->stream = sim->own_stream;
        return CHKN_OK;
    });
}

int chkn_set_launch_opts(chkn_sim* sim, const chkn_launch_opts* o)
{
    if (!sim || !o) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    chkn_launch_opts n = *o;
    if (n.tile_i == 0) n.tile_i = 32;
    if (n.tile_j == 0) n.tile_j = 8;
    if (!tile_shape_supported(n.tile_i, n.tile_j))
        return set_error(sim, CHKN_ERR_ARG, "unsupported tile shape");
    if (n.vertex_frames != 0 && n.vertex_frames != 1)
        return set_error(sim, CHKN_ERR_ARG, "vertex_frames must be 0 or 1");
    if (n.vertex_frames && !vertex_frames_supported(n.tile_i, n.tile_j, sim->x_order))
        return set_error(sim, CHKN_ERR_ARG, "vertex-frame stage kernel not built for this tile shape");
    if (n.tile_k < 0) return set_error(sim, CHKN_ERR_ARG, "tile_k must be >= 0");
    if (n.tile_order != 0 && n.tile_order != 1) return set_error(sim, CHKN_ERR_ARG, "tile_order must be 0 or 1");
    return guarded(sim, [&] {
        sim->opts = n;
        build_tiles(sim);
        return CHKN_OK;
    });
}

int chkn_get_report(chkn_sim* sim, chkn_report* rep)
{
    if (!sim || !rep) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    fill_report(sim, 0, rep, 0);
    return CHKN_OK;
}

int chkn_num_instances(chkn_sim* sim) { return sim ? (int)sim->inst.size() : -1; }

int chkn_instance_blocks(chkn_sim* sim, int inst, int* first, int* count)
{
    if (!sim || inst < 0 || inst >= (int)sim->inst.size()) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    if (first) *first = sim->inst[inst].first_block;
    if (count) *count = sim->inst[inst].nblocks;
    return CHKN_OK;
}

int chkn_get_instance_report(chkn_sim* sim, int inst, chkn_report* rep)
{
    if (!sim || !rep || inst < 0 || inst >= (int)sim->inst.size()) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    fill_report(sim, inst, rep, 0);
    return CHKN_OK;
}

int chkn_num_blocks(chkn_sim* sim) { return sim ? (int)sim->blocks.size() : -1; }

int chkn_block_dims(chkn_sim* sim, int blk, int* dims3)
{
    if (!sim || !dims3 || blk < 0 || blk >= (int)sim->blocks.size()) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    dims3[0] = sim->blocks[blk].nic; dims3[1] = sim->blocks[blk].njc; dims3[2] = sim->blocks[blk].nkc;
    return CHKN_OK;
}

long long chkn_active_cells(chkn_sim* sim)
{
    if (!sim) return -1;
    long long n = 0;
    for (auto& b : sim->blocks) if (b.active) n += b.n;
    return n;
}

int chkn_last_timing(chkn_sim* sim, double* ms, long long* launches)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    if (ms) *ms = sim->last_ms;
    if (launches) *launches = sim->last_launches;
    return CHKN_OK;
}

const char* chkn_last_error(chkn_sim* sim) { return sim ? sim->err.c_str() : g_create_error.c_str(); }

int chkn_real_size(void) { return (int)sizeof(real); }

void chkn_destroy(chkn_sim* sim)
{
    if (!sim) return;
    free_device(sim);
    delete sim;
}

}  // extern "C"
