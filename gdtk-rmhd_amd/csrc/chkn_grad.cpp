// chkn_grad.cpp -- gradients of an unrolled chicken march (reverse mode), C-ABI part.
//
// chkn_grad_forward runs n steps from the handle's current state, keeping a checkpoint (U
// and W) every C steps and the time step of every step.  chkn_grad_backward takes the
// adjoint of the final conserved state and walks the checkpoint segments backwards: each
// segment's 3C+1 stage states are read from a tape (a block table whose W points at the
// tape, so the stage kernels are the forward ones) and the steps of the segment are
// reversed with the adjoint kernels.  The last segments, as many as the tape budget holds
// in device memory, are taped by the forward pass itself (resident tapes); the others are
// recomputed from their checkpoints with the recorded time steps, the next one on a second
// stream while the current one is reversed.  The time-step schedule is held fixed (the
// gradient of the CFL min-reduction is not taken), so the gradient is that of the map with
// the recorded dt sequence.  Recomputation and taping repeat the forward arithmetic
// exactly, so the gradient depends neither on C nor on the budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/chkn.h"
#include "chkn_internal.h"
#include "chkn_runtime.h"

using namespace chkn;
using namespace chkn::rt;

namespace {

// Work tables of the adjoint kernels (faces, cells, inflow lines), built once per handle.
void build_adj_tables(chkn_sim* s)
{
    GradState& g = *s->grad;
    if (g.d_face_chunks) return;
    std::vector<AdjPatch> fc, fck;
    std::vector<CellChunk> cc;
    std::vector<InflowCol> ic;
    g.fused_ok = true;
    for (size_t ib = 0; ib < s->blocks.size(); ++ib) {
        const HostBlock& b = s->blocks[ib];
        if (!b.active) continue;
        if (s->bdev[ib].rp[4] >= 0 || s->bdev[ib].rp[5] >= 0) g.fused_ok = false;
        // Face patches per direction: lines of kAdjPatchT first transverse indices, each
        // covering kAdjPatchOwn cells along the line; i and j patches of every block first.
        for (int d = 0; d < 3; ++d) {
            const int nline = d == 0 ? b.nic : (d == 1 ? b.njc : b.nkc);
            const int na = d == 0 ? b.njc : b.nic;
            const int nc = d == 2 ? b.njc : b.nkc;
            for (int c = 0; c < nc; ++c)
                for (int a0 = 0; a0 < na; a0 += kAdjPatchT)
                    for (int m0 = 0; m0 < nline; m0 += kAdjPatchOwn)
                        (d == 2 ? fck : fc).push_back(AdjPatch{(int)ib, d, m0, a0, c});
        }
        for (int c0 = 0; c0 < b.n; c0 += 256) cc.push_back(CellChunk{(int)ib, c0, std::min(c0 + 256, b.n)});
        BlockDev& bd = s->bdev[ib];
        bd.icol = (int)ic.size();
        for (int side = 0; side < 6; ++side) {
            bd.icol_side[side] = -1;
            if (b.bc[side] != CHKN_BC_INFLOW) continue;
            bd.icol_side[side] = (int)ic.size() - bd.icol;
            const int d = side >> 1;
            const int nt = d == 0 ? b.njc * b.nkc : (d == 1 ? b.nic * b.nkc : b.nic * b.njc);
            for (int t = 0; t < nt; ++t) {
                ic.push_back(InflowCol{(int)ib, side, t});
                g.inflow_col_state.push_back(b.bc_fs[side]);
            }
        }
    }
    g.nface_ij = (int)fc.size();
    fc.insert(fc.end(), fck.begin(), fck.end());
    g.nface_chunks = (int)fc.size();
    g.ncell_chunks = (int)cc.size();
    g.ninflow_cols = (int)ic.size();
    HIPCHK(hipMalloc(&g.d_face_chunks, sizeof(AdjPatch) * std::max<size_t>(fc.size(), 1)));
    HIPCHK(hipMalloc(&g.d_cell_chunks, sizeof(CellChunk) * std::max<size_t>(cc.size(), 1)));
    HIPCHK(hipMalloc(&g.d_inflow_cols, sizeof(InflowCol) * std::max<size_t>(ic.size(), 1)));
    HIPCHK(hipMalloc(&g.d_inflow_acc, sizeof(double) * 6 * std::max<size_t>(ic.size(), 1)));
    if (!fc.empty())
        HIPCHK(hipMemcpy(g.d_face_chunks, fc.data(), sizeof(AdjPatch) * fc.size(), hipMemcpyHostToDevice));
    if (!cc.empty())
        HIPCHK(hipMemcpy(g.d_cell_chunks, cc.data(), sizeof(CellChunk) * cc.size(), hipMemcpyHostToDevice));
    if (!ic.empty())
        HIPCHK(hipMemcpy(g.d_inflow_cols, ic.data(), sizeof(InflowCol) * ic.size(), hipMemcpyHostToDevice));
}

GradState& grad_state(chkn_sim* s)
{
    if (!s->grad) s->grad.reset(new GradState());
    return *s->grad;
}

// Checkpoint slot j of every active block: U, then W (slot 0 of the ring).
void save_checkpoint(chkn_sim* s, int j)
{
    for (auto& b : s->blocks) {
        if (!b.active) continue;
        const size_t n6 = 6 * (size_t)b.n;
        double* dst = b.ckpt + (size_t)j * 2 * n6;
        HIPCHK(hipMemcpyAsync(dst, b.U, sizeof(double) * n6, hipMemcpyDeviceToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(dst + n6, b.W, sizeof(double) * n6, hipMemcpyDeviceToDevice, s->stream));
    }
}

// Restores U from checkpoint j and writes its W into `wdst(block)`, on stream st.
template <typename F>
void load_checkpoint(chkn_sim* s, int j, F wdst, hipStream_t st)
{
    for (auto& b : s->blocks) {
        if (!b.active) continue;
        const size_t n6 = 6 * (size_t)b.n;
        const double* src = b.ckpt + (size_t)j * 2 * n6;
        HIPCHK(hipMemcpyAsync(b.U, src, sizeof(double) * n6, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(wdst(b), src + n6, sizeof(double) * n6, hipMemcpyDeviceToDevice, st));
    }
}

void alloc_checkpoints(chkn_sim* s, int nckpt)
{
    GradState& g = *s->grad;
    if (g.nckpt_alloc >= nckpt) return;
    for (auto& b : s->blocks) {
        if (!b.active) continue;
        dfree(b.ckpt);
        b.ckpt = nullptr;
        HIPCHK(hipMalloc(&b.ckpt, sizeof(double) * 12 * (size_t)b.n * nckpt));
    }
    g.nckpt_alloc = nckpt;
}

// W slots (doubles) of one segment's tape of block b, and bytes of one segment's tapes over
// the active blocks.
// A tape holds every stage state of its C steps and the state after them: stages_per_step * C
// + 1 W slots.
size_t tape_doubles(const chkn_sim* s, const HostBlock& b, int C)
{
    return 6 * (size_t)b.n * ((size_t)stages_per_step(s) * C + 1);
}
size_t segment_tape_bytes(const chkn_sim* s, int C)
{
    size_t n = 0;
    for (auto& b : s->blocks)
        if (b.active) n += sizeof(double) * tape_doubles(s, b, C);
    return n;
}

void free_resident(chkn_sim* s)
{
    GradState& g = *s->grad;
    for (auto& b : s->blocks) {
        dfree(b.rtape);
        b.rtape = nullptr;
    }
    for (BlockDev* t : g.d_rtab) dfree(t);
    g.d_rtab.clear();
    g.nres_alloc = g.rtape_C = g.nres = 0;
}

// Resident tapes for the last segments of an nck-segment unroll, as many as the budget holds
// (held between unrolls while the segment length and the budget allow).
void plan_resident(chkn_sim* s, int C, int nck)
{
    GradState& g = *s->grad;
    if (g.rtape_C != C) free_resident(s);
    const size_t seg = segment_tape_bytes(s, C);
    size_t budget = 0;
    if (g.tape_budget < 0) {
        // Half of the memory free or held by the unroll's tapes (resident and recompute ring),
        // so that the budget does not shrink from one unroll to the next.
        size_t fr = 0, tot = 0, ring = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        if (g.tape_C > 0)
            for (auto& b : s->blocks)
                if (b.active) ring += sizeof(double) * tape_doubles(s, b, g.tape_C) * ((b.tape ? 1 : 0) + (b.tape2 ? 1 : 0));
        budget = (fr + (size_t)g.nres_alloc * seg + ring) / 2;
    } else {
        budget = (size_t)g.tape_budget;
    }
    int want = seg > 0 ? (int)std::min<size_t>((size_t)nck, budget / seg) : 0;
    if (want > g.nres_alloc || (g.tape_budget >= 0 && want < g.nres_alloc)) {
        free_resident(s);
        // A budget raised on a handle that holds recompute tapes: those are released (and
        // allocated again, as far as they are still needed) when the resident tapes do not fit
        // beside them.
        for (int attempt = 0; attempt < 2 && want > 0; ++attempt) {
            bool ok = true;
            for (auto& b : s->blocks) {
                if (!b.active) continue;
                if (hipMalloc(&b.rtape, sizeof(double) * tape_doubles(s, b, C) * want) != hipSuccess) {
                    (void)hipGetLastError();
                    ok = false;
                    break;
                }
            }
            if (ok) break;
            free_resident(s);
            if (attempt == 0 && g.tape_C > 0) {
                for (auto& b : s->blocks) {
                    dfree(b.tape);
                    dfree(b.tape2);
                    b.tape = b.tape2 = nullptr;
                }
                g.tape_C = 0;
                g.two_tapes = false;
            } else {
                want = 0;
            }
        }
        if (want > 0) {
            g.d_rtab.assign(want, nullptr);
            for (BlockDev*& t : g.d_rtab) HIPCHK(hipMalloc(&t, sizeof(BlockDev) * s->bdev.size()));
            g.nres_alloc = want;
            g.rtape_C = C;
        }
    }
    g.nres = std::min(want, g.nres_alloc);
}

// W of resident tape r of block b.
double* resident_tape(const chkn_sim* s, const HostBlock& b, int C, int r)
{
    return b.rtape + (size_t)r * tape_doubles(s, b, C);
}

// Block tables with W at each resident tape, with the current adjoint pointers.
void upload_resident_tables(chkn_sim* s)
{
    GradState& g = *s->grad;
    for (int r = 0; r < g.nres; ++r) {
        std::vector<BlockDev> tab = s->bdev;
        for (size_t ib = 0; ib < s->blocks.size(); ++ib)
            if (s->blocks[ib].active) tab[ib].W = resident_tape(s, s->blocks[ib], g.rtape_C, r);
        HIPCHK(hipMemcpy(g.d_rtab[r], tab.data(), sizeof(BlockDev) * tab.size(), hipMemcpyHostToDevice));
    }
}

// While alive, the handle's steps write their stage states into resident tape r: W of the
// current state is copied into its slot 0 first, and finish(K) copies the final W of a
// K-step segment back into the ring.
struct Taping {
    chkn_sim* s;
    int r, C;
    Taping(chkn_sim* sim, int res, int c) : s(sim), r(res), C(c)
    {
        for (auto& b : s->blocks)
            if (b.active)
                HIPCHK(hipMemcpyAsync(resident_tape(s, b, C, r), b.W, sizeof(double) * 6 * (size_t)b.n,
                                      hipMemcpyDeviceToDevice, s->stream));
        s->tape_tab = s->grad->d_rtab[r];
        s->tape_base = 0;
    }
    void finish(int K)
    {
        const size_t last = (size_t)stages_per_step(s) * K;
        for (auto& b : s->blocks)
            if (b.active)
                HIPCHK(hipMemcpyAsync(b.W, resident_tape(s, b, C, r) + 6 * (size_t)b.n * last,
                                      sizeof(double) * 6 * (size_t)b.n, hipMemcpyDeviceToDevice, s->stream));
    }
    ~Taping()
    {
        s->tape_tab = nullptr;
        s->tape_base = 0;
    }
};

void alloc_adjoint(chkn_sim* s, int C, bool ring)
{
    GradState& g = *s->grad;
    if (!ring && g.tape_C > 0) {  // every segment is on a resident tape: no recompute ring
        for (auto& b : s->blocks) {
            dfree(b.tape);
            dfree(b.tape2);
            b.tape = b.tape2 = nullptr;
        }
        g.tape_C = 0;
        g.two_tapes = false;
    }
    for (size_t ib = 0; ib < s->blocks.size(); ++ib) {
        HostBlock& b = s->blocks[ib];
        if (!b.active) continue;
        if (ring && g.tape_C < C) {
            dfree(b.tape);
            dfree(b.tape2);
            b.tape = b.tape2 = nullptr;
            HIPCHK(hipMalloc(&b.tape, sizeof(double) * tape_doubles(s, b, C)));
        }
        if (!b.adj) {
            // Ghost lines of the six sides, two depths of six values each.
            const size_t lines = 2 * ((size_t)b.njc * b.nkc + (size_t)b.nic * b.nkc + (size_t)b.nic * b.njc);
            HIPCHK(hipMalloc(&b.adj, sizeof(double) * 24 * (size_t)b.n));
            HIPCHK(hipMalloc(&b.Gc, sizeof(double) * 18 * (size_t)b.n));
            HIPCHK(hipMalloc(&b.GB, sizeof(double) * 12 * lines));
            s->bdev[ib].adj = b.adj;
            s->bdev[ib].Gc = b.Gc;
            s->bdev[ib].GB = b.GB;
        }
    }
    if (ring && g.tape_C < C) {
        // The second tape is optional: without it the segments are recomputed and reversed
        // one after the other.
        g.two_tapes = true;
        for (auto& b : s->blocks) {
            if (!b.active) continue;
            if (hipMalloc(&b.tape2, sizeof(double) * tape_doubles(s, b, C)) != hipSuccess) {
                (void)hipGetLastError();
                b.tape2 = nullptr;
                g.two_tapes = false;
                break;
            }
        }
        if (!g.two_tapes)
            for (auto& b : s->blocks) {
                dfree(b.tape2);
                b.tape2 = nullptr;
            }
    }
    if (ring) g.tape_C = std::max(g.tape_C, C);
    if (!g.rec) {
        HIPCHK(hipStreamCreateWithFlags(&g.rec, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g.ev_rev[0], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&g.ev_rev[1], hipEventDisableTiming));
        HIPCHK(hipMalloc(&g.d_tab[0], sizeof(BlockDev) * s->bdev.size()));
        HIPCHK(hipMalloc(&g.d_tab[1], sizeof(BlockDev) * s->bdev.size()));
    }
}

// Block tables with W at tape (t = 0) and tape2 (t = 1) and at the resident tapes, with the
// current adjoint pointers.
void upload_tape_tables(chkn_sim* s)
{
    GradState& g = *s->grad;
    upload_resident_tables(s);
    for (int t = 0; t < 2; ++t) {
        std::vector<BlockDev> tab = s->bdev;
        for (size_t ib = 0; ib < s->blocks.size(); ++ib)
            if (s->blocks[ib].active) tab[ib].W = t == 0 ? s->blocks[ib].tape : s->blocks[ib].tape2;
        HIPCHK(hipMemcpy(g.d_tab[t], tab.data(), sizeof(BlockDev) * tab.size(), hipMemcpyHostToDevice));
    }
}

// Queues the recomputation of checkpoint segment j (K steps) into tape t on stream st:
// checkpoint load, control area, the K steps' stage launches, control area read-back.  The caller
// synchronises st before reading the bad-cell counters (recompute_bad).
void queue_recompute(chkn_sim* s, int j, int K, int t, hipStream_t st)
{
    GradState& g = *s->grad;
    const int N = (int)s->inst.size();
    load_checkpoint(s, j, [t](HostBlock& b) { return t == 0 ? b.tape : b.tape2; }, st);
    const double* dts = g.dts.data() + (size_t)j * g.C * N;
    for (int q = 0; q < K; ++q) {
        double* dtv = ctl_dt(s->h_ctl, q, N);
        int* run = ctl_run(s->h_ctl, q, N);
        unsigned long long* sg = ctl_sig(s->h_ctl, q, N);
        int* bad = ctl_bad(s->h_ctl, q, N);
        for (int qi = 0; qi < N; ++qi) {
            dtv[qi] = dts[(size_t)q * N + qi];
            run[qi] = 1;
            sg[qi] = ~0ull;
        }
        bad[0] = bad[1] = bad[2] = 0;
        if (s->any_reacting) fill_chem(s, q);
    }
    HIPCHK(hipMemcpyAsync(s->d_ctl, s->h_ctl, s->ctl_bytes, hipMemcpyHostToDevice, st));
    for (int q = 0; q < K; ++q) {
        StageArgs a = stage_args(s, q);
        a.blocks = g.d_tab[t];
        a.tile_base = 0;
        a.ntiles = s->ntiles;
        const int nst = stages_per_step(s);
        for (int pos = 1; pos <= nst; ++pos) {
            a.win = nst * q + pos - 1;
            a.wout = nst * q + pos;
            launch_stage_form(s, a, (s->time_scheme == CHKN_TIME_PC && pos == 2) ? 4 : pos, false, false, st);
        }
        // The step's reaction on its last stage state (as the forward step, simulate.cu:296-303).
        if (s->any_reacting)
            HIPCHK(launch_chemistry(a.blocks, s->d_split_cells, s->nsplit_cells, a.ctl, a.wout, false, st));
    }
    HIPCHK(hipMemcpyAsync(s->h_ctl, s->d_ctl, s->ctl_bytes, hipMemcpyDeviceToHost, st));
}

int recompute_bad(chkn_sim* s, int K)
{
    const int N = (int)s->inst.size();
    for (int q = 0; q < K; ++q) {
        const int* bad = ctl_bad(s->h_ctl, q, N);
        for (int st = 0; st < 3; ++st)
            if (bad[st] > 0) return bad[st];
    }
    return 0;
}

// K steps with the time steps dts[q*N + inst] from the handle's current state, on the
// stepping ring.  Returns the first stage's bad-cell count (0 if none).
int fixed_dt_steps(chkn_sim* s, int K, const double* dts)
{
    const int N = (int)s->inst.size();
    for (int q = 0; q < K; ++q) {
        double* dtv = ctl_dt(s->h_ctl, q, N);
        int* run = ctl_run(s->h_ctl, q, N);
        unsigned long long* sg = ctl_sig(s->h_ctl, q, N);
        int* bad = ctl_bad(s->h_ctl, q, N);
        for (int qi = 0; qi < N; ++qi) {
            dtv[qi] = dts[(size_t)q * N + qi];
            run[qi] = 1;
            sg[qi] = ~0ull;
        }
        bad[0] = bad[1] = bad[2] = 0;
        if (s->any_reacting) fill_chem(s, q);
    }
    run_segment(s, K, false);
    for (int q = 0; q < K; ++q) {
        const int* bad = ctl_bad(s->h_ctl, q, N);
        for (int st = 0; st < 3; ++st)
            if (bad[st] > 0) return bad[st];
    }
    return 0;
}

// The unroll length is explicit: the instances' max_step and max_time do not stop it.
struct LiftLimits {
    chkn_sim* s;
    std::vector<std::pair<int, double>> saved;
    explicit LiftLimits(chkn_sim* sim) : s(sim)
    {
        for (auto& I : s->inst) {
            saved.emplace_back(I.max_step, I.max_time);
            I.max_step = 0x7fffffff;
            I.max_time = HUGE_VAL;
        }
    }
    ~LiftLimits()
    {
        for (size_t q = 0; q < s->inst.size(); ++q) {
            s->inst[q].max_step = saved[q].first;
            s->inst[q].max_time = saved[q].second;
        }
    }
};

int grad_forward(chkn_sim* s, int nsteps, int C, const double* dt_in, double* dt_out)
{
    if (s->failed) return set_error(s, CHKN_ERR_STATE, "simulation is in a failed state");
    if (nsteps <= 0) return set_error(s, CHKN_ERR_ARG, "nsteps must be positive");
    if (C <= 0 || C > kMaxSeg) return set_error(s, CHKN_ERR_ARG, "checkpoint interval must be in 1..64");
    if (s->viscous) return set_error(s, CHKN_ERR_ARG, "the adjoint of the viscous fluxes is not built");
    // The unroll is a function of the conserved state: gas states follow from it, also in
    // the first stage (the recomputation from checkpoints does the same).
    s->fresh = false;
    LiftLimits lift(s);
    GradState& g = grad_state(s);
    const int N = (int)s->inst.size();
    const int nck = (nsteps + C - 1) / C;
    alloc_checkpoints(s, nck + 1);  // + 1 slot to park the final state during the backward pass
    g.valid = false;
    plan_resident(s, C, nck);
    upload_resident_tables(s);
    g.C = C;
    g.nsteps = nsteps;
    g.dts.assign((size_t)nsteps * N, 0.0);
    s->last_ms = 0.0;
    s->last_launches = 0;
    double ms_total = 0.0;
    long long launches = 0;
    for (int j = 0; j < nck; ++j) {
        const int K = std::min(C, nsteps - j * C);
        save_checkpoint(s, j);
        std::unique_ptr<Taping> tp;
        if (j >= nck - g.nres) tp.reset(new Taping(s, j - (nck - g.nres), C));
        if (dt_in) {
            const double* dts = dt_in + (size_t)j * C * N;
            const int bad = fixed_dt_steps(s, K, dts);
            ms_total += s->last_ms;
            launches += s->last_launches;
            s->last_ms = 0.0;
            s->last_launches = 0;
            if (bad > 0) {
                s->failed = true;
                return set_error(s, bad, "bad cells in the gradient forward pass");
            }
            for (int qi = 0; qi < N; ++qi) {
                Instance& I = s->inst[qi];
                for (int q = 0; q < K; ++q) {
                    I.dt = dts[(size_t)q * N + qi];
                    I.t += I.dt;
                }
                I.step += K;
                I.steps_since_last_plot += K;
                I.have_sig = false;
            }
            s->gstep += K;
            s->fresh = false;
            std::copy(dts, dts + (size_t)K * N, g.dts.begin() + (size_t)j * C * N);
        } else {
            std::vector<double> log;
            s->dt_log = &log;
            const int rc = do_steps(s, K, false, true, nullptr);
            s->dt_log = nullptr;
            ms_total += s->last_ms;
            launches += s->last_launches;
            if (rc != CHKN_OK) return rc;
            if ((int)log.size() != K * N) return set_error(s, CHKN_ERR_STATE, "an instance stopped during the unroll");
            for (double v : log)
                if (v == 0.0) return set_error(s, CHKN_ERR_STATE, "an instance stopped during the unroll");
            std::copy(log.begin(), log.end(), g.dts.begin() + (size_t)j * C * N);
        }
        if (tp) tp->finish(K);
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->last_ms = ms_total;
    s->last_launches = launches;
    if (dt_out) std::copy(g.dts.begin(), g.dts.end(), dt_out);
    g.valid = true;
    g.generation += 1;
    return CHKN_OK;
}

size_t flat_size(chkn_sim* s)
{
    size_t n = 0;
    for (auto& b : s->blocks) n += 6 * (size_t)b.n;
    return n;
}

int grad_backward(chkn_sim* s, const double* lam, double* grad_u0, double* grad_inflow)
{
    GradState& g = grad_state(s);
    if (!g.valid) return set_error(s, CHKN_ERR_STATE, "chkn_grad_backward needs a chkn_grad_forward first");
    const int N = (int)s->inst.size();
    const int C = g.C, nsteps = g.nsteps;
    const int nck = (nsteps + C - 1) / C;
    const int nres = g.nres;
    const int first_res = nck - nres;  // segments first_res .. nck-1 are on resident tapes
    alloc_adjoint(s, C, first_res > 0);
    build_adj_tables(s);
    dfree(g.d_dts);
    g.d_dts = nullptr;
    HIPCHK(hipMalloc(&g.d_dts, sizeof(double) * g.dts.size()));
    HIPCHK(hipMemcpy(g.d_dts, g.dts.data(), sizeof(double) * g.dts.size(), hipMemcpyHostToDevice));
    dfree(g.d_chem);
    g.d_chem = nullptr;
    if (s->any_reacting) {
        // exp(-alpha dt) of every step, computed as the forward step computed it (fill_chem).
        std::vector<double> chem(g.dts.size());
        for (size_t q = 0; q < chem.size(); ++q)
            chem[q] = s->inst[q % N].reacting ? (double)std::exp(-kAlpha * (real)g.dts[q]) : -1.0;
        HIPCHK(hipMalloc(&g.d_chem, sizeof(double) * chem.size()));
        HIPCHK(hipMemcpy(g.d_chem, chem.data(), sizeof(double) * chem.size(), hipMemcpyHostToDevice));
    }
    hipEvent_t e0 = s->ev0, e1 = s->ev1;
    HIPCHK(hipEventRecord(e0, s->stream));
    // Park the final state; seed A0 with the final adjoint.
    save_checkpoint(s, nck);
    size_t off = 0;
    for (auto& b : s->blocks) {
        const size_t n6 = 6 * (size_t)b.n;
        if (b.active)
            HIPCHK(hipMemcpyAsync(b.adj, lam + off, sizeof(double) * n6, hipMemcpyDeviceToDevice, s->stream));
        off += n6;
    }
    if (g.ninflow_cols > 0)
        HIPCHK(hipMemsetAsync(g.d_inflow_acc, 0, sizeof(double) * 6 * g.ninflow_cols, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    upload_tape_tables(s);
    AdjArgs A;
    std::memset(&A, 0, sizeof A);
    A.inflow = s->d_inflow;
    A.patches = (const AdjPatch*)g.d_face_chunks;
    A.cells = (const CellChunk*)g.d_cell_chunks;
    A.icols = (const InflowCol*)g.d_inflow_cols;
    A.inflow_acc = g.d_inflow_acc;
    A.nicols = g.ninflow_cols;
    // Seeds of a reversed step: the weights of the stages' residuals in U^{n+1}.
    A.seed_p = s->time_scheme == CHKN_TIME_PC ? 0.5 : kOneSixth;
    A.seed_d = s->time_scheme == CHKN_TIME_PC ? 0.5 : kFourSixth;
    A.seed_p_stored = s->time_scheme == CHKN_TIME_PC ? 1 : 0;
    // D is double-buffered in the fused form (the k patches finish cells whose D neighbouring
    // patches still read): dcur is the buffer the next face pass reads, which seeds write.
    const bool fused = g.fused_ok && !g.separate_gather;
    int dcur = 12;
    A.dr = A.dw = dcur;
    const AdjPatch* kpatches = (const AdjPatch*)g.d_face_chunks + g.nface_ij;
    long long launches = 0;
    int rc = CHKN_OK;
    // Segments below first_res are recomputed.  With two tapes, segment j-1 is recomputed on
    // the second stream while segment j is reversed (the first one while the resident
    // segments are); the recompute into a tape waits for the reversal that last read it.
    const bool two = g.two_tapes && first_res > 0;
    hipStream_t rs = two ? g.rec : s->stream;
    auto seg_len = [&](int j) { return std::min(C, nsteps - j * C); };
    if (first_res > 0) queue_recompute(s, first_res - 1, seg_len(first_res - 1), two ? ((first_res - 1) & 1) : 0, rs);
    for (int j = nck - 1; j >= 0; --j) {
        const int K = seg_len(j);
        const int t = two ? (j & 1) : 0;
        const bool resident = j >= first_res;
        if (resident) {
            A.blocks = g.d_rtab[j - first_res];
        } else {
            HIPCHK(hipStreamSynchronize(rs));
            launches += (long long)stages_per_step(s) * K;
            if (recompute_bad(s, K) > 0) {
                rc = set_error(s, CHKN_ERR_STATE, "bad cells while recomputing a checkpoint segment");
                break;
            }
            A.blocks = g.d_tab[t];
        }
        for (int q = K - 1; q >= 0; --q) {
            const int step = j * C + q;
            A.dt = g.d_dts + (size_t)step * N;
            const int nst = stages_per_step(s);
            A.dr = A.dw = dcur;
            if (s->any_reacting) {
                // The step's reaction first (its state after the reaction is the next step's
                // first slot), then the seeds; the stage-1 gathers do not seed.
                A.dt_next = nullptr;
                A.chem = g.d_chem + (size_t)step * N;
                A.slot = nst * (q + 1);
                HIPCHK(launch_adj_chem_seed(A, g.ncell_chunks, s->stream));
                launches += 1;
            } else {
                A.dt_next = step > 0 ? g.d_dts + (size_t)(step - 1) * N : nullptr;
                if (step == nsteps - 1) {
                    HIPCHK(launch_adj_seed(A, g.ncell_chunks, s->stream));
                    launches += 1;
                }
            }
            // Stage positions in reverse; the predictor-corrector's corrector (position 2)
            // gathers as the TVD-RK3 stage 2 (D = P + dt Ubar), its predictor as stage 1.
            for (int pos = nst; pos >= 1; --pos) {
                A.slot = nst * q + pos - 1;
                A.dr = dcur;
                if (fused) {
                    A.dw = 30 - dcur;  // 12 <-> 18
                    HIPCHK(launch_adj_faces(A, g.nface_ij, s->x_order, s->flux_calc, 0, s->stream));
                    AdjArgs Ak = A;
                    Ak.patches = kpatches;
                    HIPCHK(launch_adj_faces(Ak, g.nface_chunks - g.nface_ij, s->x_order, s->flux_calc, pos,
                                            s->stream));
                } else {
                    A.dw = dcur;
                    HIPCHK(launch_adj_faces(A, g.nface_chunks, s->x_order, s->flux_calc, 0, s->stream));
                    HIPCHK(launch_adj_gather(A, g.ncell_chunks, pos, s->stream));
                }
                dcur = A.dw;
                launches += 2;
            }
        }
        if (!resident && j > 0) {
            if (two) {
                HIPCHK(hipEventRecord(g.ev_rev[t], s->stream));
                // Tape t^1 was last read by the reversal of segment j+1 (if that was recomputed).
                if (j + 1 < first_res) HIPCHK(hipStreamWaitEvent(rs, g.ev_rev[t ^ 1], 0));
            }
            queue_recompute(s, j - 1, seg_len(j - 1), two ? (t ^ 1) : 0, rs);
        }
    }
    HIPCHK(hipStreamSynchronize(rs));
    HIPCHK(hipStreamSynchronize(s->stream));
    // Restore the final state into the stepping ring.
    load_checkpoint(s, nck, [](HostBlock& b) { return b.W; }, s->stream);
    if (rc == CHKN_OK) {
        off = 0;
        for (auto& b : s->blocks) {
            const size_t n6 = 6 * (size_t)b.n;
            if (b.active)
                HIPCHK(hipMemcpyAsync(grad_u0 + off, b.adj, sizeof(double) * n6, hipMemcpyDeviceToDevice, s->stream));
            else
                HIPCHK(hipMemsetAsync(grad_u0 + off, 0, sizeof(double) * n6, s->stream));
            off += n6;
        }
    }
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    s->last_ms = ms;
    s->last_launches = launches;
    if (rc != CHKN_OK) return rc;
    if (grad_inflow) {
        const int ns = (int)s->inflow_w.size() / 6;
        std::fill(grad_inflow, grad_inflow + 6 * (size_t)ns, 0.0);
        std::vector<double> acc(6 * (size_t)g.ninflow_cols);
        if (g.ninflow_cols > 0)
            HIPCHK(hipMemcpy(acc.data(), g.d_inflow_acc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < g.ninflow_cols; ++c)
            for (int q = 0; q < 6; ++q)
                grad_inflow[6 * (size_t)g.inflow_col_state[c] + q] += acc[(size_t)q * g.ninflow_cols + c];
    }
    return CHKN_OK;
}

}  // namespace

namespace chkn {
namespace rt {

void free_grad(chkn_sim* s)
{
    for (auto& b : s->blocks) {
        dfree(b.tape); dfree(b.tape2); dfree(b.ckpt); dfree(b.adj); dfree(b.Gc); dfree(b.GB);
        b.tape = b.tape2 = b.ckpt = b.adj = b.Gc = b.GB = nullptr;
    }
    if (s->grad) {
        free_resident(s);
        dfree(s->grad->d_face_chunks); dfree(s->grad->d_cell_chunks);
        dfree(s->grad->d_inflow_cols); dfree(s->grad->d_inflow_acc); dfree(s->grad->d_dts); dfree(s->grad->d_chem);
        dfree(s->grad->d_tab[0]); dfree(s->grad->d_tab[1]);
        if (s->grad->ev_rev[0]) (void)hipEventDestroy(s->grad->ev_rev[0]);
        if (s->grad->ev_rev[1]) (void)hipEventDestroy(s->grad->ev_rev[1]);
        if (s->grad->rec) (void)hipStreamDestroy(s->grad->rec);
        s->grad.reset();
    }
}

}  // namespace rt
}  // namespace chkn

extern "C" {

int chkn_grad_forward(chkn_sim* sim, int nsteps, int checkpoint_every, const double* dt_in,
                      double* dt_out)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    return guarded(sim, [&] { return grad_forward(sim, nsteps, checkpoint_every, dt_in, dt_out); });
}

int chkn_grad_backward(chkn_sim* sim, const double* lambda_dev, double* grad_u0_dev,
                       double* grad_inflow)
{
    if (!sim || !lambda_dev || !grad_u0_dev) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] { return grad_backward(sim, lambda_dev, grad_u0_dev, grad_inflow); });
}

long long chkn_state_size(chkn_sim* sim) { return sim ? (long long)flat_size(sim) : -1; }

int chkn_grad_set_tape_budget(chkn_sim* sim, long long bytes)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    return guarded(sim, [&] {
        grad_state(sim).tape_budget = bytes < 0 ? -1 : bytes;
        return CHKN_OK;
    });
}

int chkn_grad_set_gather_form(chkn_sim* sim, int separate)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    return guarded(sim, [&] {
        grad_state(sim).separate_gather = separate != 0;
        return CHKN_OK;
    });
}

int chkn_grad_gather_form(chkn_sim* sim)
{
    if (!sim) return set_error(nullptr, CHKN_ERR_ARG, "null handle");
    if (sim->grad && sim->grad->separate_gather) return 1;
    for (size_t ib = 0; ib < sim->blocks.size(); ++ib)
        if (sim->blocks[ib].active && (sim->bdev[ib].rp[4] >= 0 || sim->bdev[ib].rp[5] >= 0)) return 1;
    return 0;
}

int chkn_grad_resident_segments(chkn_sim* sim)
{
    if (!sim) return -1;
    return sim->grad ? sim->grad->nres : 0;
}

long long chkn_grad_generation(chkn_sim* sim)
{
    if (!sim) return -1;
    return sim->grad ? sim->grad->generation : 0;
}

int chkn_num_flow_states(chkn_sim* sim) { return sim ? (int)sim->inflow_w.size() / 6 : -1; }

int chkn_read_conserved_dev(chkn_sim* sim, double* dst_dev)
{
    if (!sim || !dst_dev) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        size_t off = 0;
        for (auto& b : sim->blocks) {
            const size_t n6 = 6 * (size_t)b.n;
            if (b.active)
                HIPCHK(hipMemcpyAsync(dst_dev + off, b.U, sizeof(double) * n6, hipMemcpyDeviceToDevice, sim->stream));
            else
                HIPCHK(hipMemsetAsync(dst_dev + off, 0, sizeof(double) * n6, sim->stream));
            off += n6;
        }
        HIPCHK(hipStreamSynchronize(sim->stream));
        return CHKN_OK;
    });
}

int chkn_write_conserved_dev(chkn_sim* sim, const double* src_dev)
{
    if (!sim || !src_dev) return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        GradState& g = grad_state(sim);
        build_adj_tables(sim);
        size_t off = 0;
        for (auto& b : sim->blocks) {
            const size_t n6 = 6 * (size_t)b.n;
            if (b.active)
                HIPCHK(hipMemcpyAsync(b.U, src_dev + off, sizeof(double) * n6, hipMemcpyDeviceToDevice, sim->stream));
            off += n6;
        }
        HIPCHK(launch_decode_all(sim->d_blocks, (const CellChunk*)g.d_cell_chunks, g.ncell_chunks, sim->stream));
        HIPCHK(hipStreamSynchronize(sim->stream));
        sim->fresh = false;
        sim->failed = false;
        for (auto& I : sim->inst) I.have_sig = false;
        g.valid = false;
        return CHKN_OK;
    });
}

int chkn_set_inflow_state(chkn_sim* sim, int state, const double* w6)
{
    if (!sim || !w6 || state < 0 || state >= (int)sim->inflow_w.size() / 6)
        return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        std::copy(w6, w6 + 6, sim->inflow_w.begin() + 6 * (size_t)state);
        HIPCHK(hipMemcpy(sim->d_inflow + 6 * (size_t)state, w6, sizeof(double) * 6, hipMemcpyHostToDevice));
        split_set_inflow(sim, state);
        if (sim->grad) sim->grad->valid = false;
        return CHKN_OK;
    });
}

int chkn_get_inflow_state(chkn_sim* sim, int state, double* w6)
{
    if (!sim || !w6 || state < 0 || state >= (int)sim->inflow_w.size() / 6)
        return set_error(sim, CHKN_ERR_ARG, "bad argument");
    std::copy(sim->inflow_w.begin() + 6 * (size_t)state, sim->inflow_w.begin() + 6 * (size_t)state + 6, w6);
    return CHKN_OK;
}

int chkn_boundary_face_areas(chkn_sim* sim, int blk, int side, double* out)
{
    if (!sim || !out || blk < 0 || blk >= (int)sim->blocks.size() || side < 0 || side > 5)
        return set_error(sim, CHKN_ERR_ARG, "bad argument");
    return guarded(sim, [&] {
        HostBlock& b = sim->blocks[blk];
        if (!b.active) return set_error(sim, CHKN_ERR_ARG, "inactive block");
        const int d = side >> 1, s = side & 1;
        const int nic = b.nic, njc = b.njc, nkc = b.nkc;
        const double* g = d == 0 ? b.gi : (d == 1 ? b.gj : b.gk);
        const size_t nf = d == 0 ? (size_t)(nic + 1) * njc * nkc
                                 : (d == 1 ? (size_t)nic * (njc + 1) * nkc : (size_t)nic * njc * (nkc + 1));
        std::vector<double> area(nf);
        HIPCHK(hipMemcpy(area.data(), g, sizeof(double) * nf, hipMemcpyDeviceToHost));
        size_t o = 0;
        if (d == 0) {
            for (int k = 0; k < nkc; ++k)
                for (int j = 0; j < njc; ++j) out[o++] = area[(s ? nic : 0) + (size_t)(nic + 1) * (j + (size_t)njc * k)];
        } else if (d == 1) {
            for (int k = 0; k < nkc; ++k)
                for (int i = 0; i < nic; ++i) out[o++] = area[i + (size_t)nic * ((s ? njc : 0) + (size_t)(njc + 1) * k)];
        } else {
            for (int j = 0; j < njc; ++j)
                for (int i = 0; i < nic; ++i) out[o++] = area[i + (size_t)nic * (j + (size_t)njc * (s ? nkc : 0))];
        }
        return CHKN_OK;
    });
}

int chkn_flux_vjp_points(int npts, int x_order, const double* in, double* out)
{
    if (npts <= 0 || !in || !out || x_order < 0 || x_order > 2) return set_error(nullptr, CHKN_ERR_ARG, "bad argument");
    double *din = nullptr, *dout = nullptr;
    try {
        HIPCHK(hipMalloc(&din, sizeof(double) * 39 * (size_t)npts));
        HIPCHK(hipMalloc(&dout, sizeof(double) * 24 * (size_t)npts));
        HIPCHK(hipMemcpy(din, in, sizeof(double) * 39 * (size_t)npts, hipMemcpyHostToDevice));
        HIPCHK(launch_flux_vjp_points(din, dout, npts, x_order, nullptr));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(out, dout, sizeof(double) * 24 * (size_t)npts, hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        dfree(din);
        dfree(dout);
        return set_error(nullptr, CHKN_ERR_HIP, e.what());
    }
    dfree(din);
    dfree(dout);
    return CHKN_OK;
}

int chkn_flux_vjp_points_noalloc(int npts, int x_order, const double* in, double* out)
{
    if (npts <= 0 || !in || !out || x_order < 0 || x_order > 2) return set_error(nullptr, CHKN_ERR_ARG, "bad argument");
    if (npts > kNoAllocPoints)
        return set_error(nullptr, CHKN_ERR_ARG,
                         "npts exceeds the allocation-free limit of " + std::to_string(kNoAllocPoints) + " points");
    const hipError_t e = flux_vjp_points_noalloc(in, out, npts, x_order);
    if (e != hipSuccess) return set_error(nullptr, CHKN_ERR_HIP, std::string("HIP error ") + hipGetErrorString(e));
    return CHKN_OK;
}

}  // extern "C"
