// chkn_split.cpp -- host side of the face-list form of a stage: which stages need it, the
// device data of the viscous gradients, the gas states the reference keeps verbatim, work
// tables and launches.
//
// The reference holds a full gas state (p, T, a besides rho and e) in every cell.  After
// a cell's first update it follows from (rho, e) through update_from_rhoe (gas.cu:77-83);
// before it, and forever in inactive blocks, it is what the flow file held
// (cell.cu:112-131); inflow ghost cells hold the configured state after update_from_pT
// (config.cu:398-406).  Only the first-order flux (p and a of the cells beside a face,
// face.cu:356-371) and the viscous gradients (T of the cloud cells, block.cu:109-131) read
// these values; the fused kernel derives them from (rho, e), so a job runs its stages in the
// face-list form when the two differ somewhere it matters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "chkn_internal.h"
#include "chkn_physics.h"

using namespace chkn;

#include "chkn_runtime.h"

namespace chkn {
namespace rt {

namespace {

constexpr int kChunk = 256;  // faces or cells per work-group

// The equation of state reproduces (p, T, a) from (rho, e, YB) bit for bit.
bool eos_agrees(real rho, real e, real YB, real p, real T, real a)
{
    const real T2 = eos_T(e, YB);
    return T2 == T && eos_p(rho, T2) == p && eos_a(T2) == a;
}

bool need_gas(const chkn_sim* s) { return s->x_order == 1 || s->viscous || s->flux_calc == 1; }

size_t face_count(const HostBlock& b)
{
    return (size_t)(b.nic + 1) * b.njc * b.nkc + (size_t)b.nic * (b.njc + 1) * b.nkc +
           (size_t)b.nic * b.njc * (b.nkc + 1);
}

// The face-list form runs every stage: viscous jobs, SBP-ASF jobs, and first-order jobs in which an inflow
// state used by an active block, or an inactive block that an active block exchanges with,
// holds gas states the equation of state does not reproduce.
bool split_always(const chkn_sim* s)
{
    if (s->viscous || s->flux_calc == 1) return true;
    if (s->x_order != 1) return false;
    const int nb = (int)s->blocks.size();
    for (int ib = 0; ib < nb; ++ib) {
        const HostBlock& b = s->blocks[ib];
        if (!b.active) continue;
        for (int f = 0; f < 6; ++f) {
            if (b.bc[f] == CHKN_BC_INFLOW_FUNCTION) return true;  // ghosts set by update_from_pT
            if (b.bc[f] == CHKN_BC_INFLOW && s->inflow_differs[b.bc_fs[f]]) return true;
            if (b.bc[f] == CHKN_BC_EXCHANGE) {
                const int o = s->bdev[ib].bc_arg[f];
                if (!s->blocks[o].active && s->blocks[o].gas_differs) return true;
            }
        }
    }
    return false;
}

}  // namespace

bool fresh_needs_file_gas(const chkn_sim* s)
{
    if (!s->fresh || !need_gas(s)) return false;
    for (const auto& b : s->blocks)
        if (b.active && b.gas_differs) return true;
    return false;
}

void split_block_setup(chkn_sim* s, HostBlock& b, BlockDev& bd, const std::vector<real>& fpos)
{
    for (int f = 0; f < 6; ++f) bd.twall[f] = b.twall[f];
    if (!b.active || !need_gas(s)) return;
    const size_t nft = face_count(b);
    dalloc(s, (void**)&b.F, sizeof(real) * 6 * nft);
    bd.F = b.F;
    if (!s->viscous) return;
    // Centres as (x, y, z, 0) records: one cloud point is two vector loads.
    auto records = [](const std::vector<real>& soa, size_t m) {
        std::vector<real> r(4 * m, 0.0);
        for (size_t q = 0; q < m; ++q)
            for (int d = 0; d < 3; ++d) r[4 * q + d] = soa[d * m + q];
        return r;
    };
    const size_t n = (size_t)b.n;
    const std::vector<real> fr = records(fpos, nft), cr = records(b.pos, n);
    dalloc(s, (void**)&b.FS, sizeof(real) * 4 * nft);
    dalloc(s, (void**)&b.fpos, sizeof(real) * 4 * nft);
    dalloc(s, (void**)&b.cpos, sizeof(real) * 4 * n);
    HIPCHK(hipMemcpy(b.fpos, fr.data(), sizeof(real) * 4 * nft, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.cpos, cr.data(), sizeof(real) * 4 * n, hipMemcpyHostToDevice));
    bd.FS = b.FS;
    bd.fpos = b.fpos;
    bd.cpos = b.cpos;
}

void side_data_setup(HostBlock& b, BlockDev& bd, const std::vector<real>& fpos, const bl::Profile* prof)
{
    if (!b.active) return;
    const int nic = b.nic, njc = b.njc, nkc = b.nkc;
    const size_t n = (size_t)b.n;
    const size_t nfi = (size_t)(nic + 1) * njc * nkc, nfj = (size_t)nic * (njc + 1) * nkc;
    const size_t nft = face_count(b);
    for (int bf = 0; bf < 6; ++bf) {
        const int dir = bf / 2, hi = bf & 1;
        const int na = dir == 0 ? njc : nic, nc = dir == 2 ? njc : nkc, nt = na * nc;
        std::vector<real> data;
        if (b.bc[bf] == CHKN_BC_INFLOW_FUNCTION) {
            data.assign((size_t)22 * nt, 0.0);
            for (int c = 0; c < nc; ++c)
                for (int a = 0; a < na; ++a) {
                    const int t = a + na * c;
                    size_t ff, cell;  // boundary face and the interior cell beside it
                    if (dir == 0) {
                        ff = (size_t)(hi ? nic : 0) + (size_t)(nic + 1) * (a + (size_t)njc * c);
                        cell = (size_t)(hi ? nic - 1 : 0) + (size_t)nic * (a + (size_t)njc * c);
                    } else if (dir == 1) {
                        ff = nfi + (size_t)a + (size_t)nic * ((hi ? njc : 0) + (size_t)(njc + 1) * c);
                        cell = (size_t)a + (size_t)nic * ((hi ? njc - 1 : 0) + (size_t)njc * c);
                    } else {
                        ff = nfi + nfj + (size_t)a + (size_t)nic * (c + (size_t)njc * (hi ? nkc : 0));
                        cell = (size_t)a + (size_t)nic * (c + (size_t)njc * (hi ? nkc - 1 : 0));
                    }
                    // Ghost positions f + d and f + 3d, d = f - c (block.cu:656-780).
                    real f[3], d[3];
                    for (int q = 0; q < 3; ++q) {
                        f[q] = fpos[q * nft + ff];
                        d[q] = f[q] - b.pos[q * n + cell];
                    }
                    const real y0 = f[1] + d[1];
                    const real y1 = f[1] + d[1] * real(3.0);  // Vector3::mul(number)
                    const bl::State st[2] = {bl::inflow_function(b.bc_fun[bf], prof, y0),
                                             bl::inflow_function(b.bc_fun[bf], prof, y1)};
                    for (int dep = 0; dep < 2; ++dep) {
                        const real v[9] = {st[dep].rho, st[dep].e, st[dep].YB, st[dep].vx, st[dep].vy,
                                             st[dep].vz, st[dep].p, st[dep].T, st[dep].a};
                        for (int q = 0; q < 9; ++q) data[(size_t)(dep * 9 + q) * nt + t] = v[q];
                    }
                    const bl::State fs = bl::inflow_function(b.bc_fun[bf], prof, f[1]);
                    data[(size_t)18 * nt + t] = fs.T;
                    data[(size_t)19 * nt + t] = fs.vx;
                    data[(size_t)20 * nt + t] = fs.vy;
                    data[(size_t)21 * nt + t] = fs.vz;
                }
        } else if (b.bc[bf] == CHKN_BC_WALL_NO_SLIP_FIXED_T && !b.twall_faces[bf].empty()) {
            data = b.twall_faces[bf];
        } else {
            continue;
        }
        HIPCHK(hipMalloc(&b.side[bf], sizeof(real) * data.size()));
        HIPCHK(hipMemcpy(b.side[bf], data.data(), sizeof(real) * data.size(), hipMemcpyHostToDevice));
        bd.side[bf] = b.side[bf];
    }
}

void split_block_gas(chkn_sim* s, HostBlock& b, BlockDev& bd)
{
    if (!need_gas(s)) return;
    const size_t n = (size_t)b.n;
    const real* io = b.io0.data();
    std::vector<real> g(3 * n);
    bool differs = false;
    for (size_t c = 0; c < n; ++c) {
        const real p = io[4 * n + c], T = io[5 * n + c], rho = io[6 * n + c], e = io[7 * n + c];
        const real YB = io[8 * n + c], a = io[9 * n + c];
        g[c] = p; g[n + c] = T; g[2 * n + c] = a;
        if (!eos_agrees(rho, e, YB, p, T, a)) differs = true;
    }
    b.gas_differs = differs;
    if (differs) {
        if (!b.gas) HIPCHK(hipMalloc(&b.gas, sizeof(real) * 3 * n));
        HIPCHK(hipMemcpy(b.gas, g.data(), sizeof(real) * 3 * n, hipMemcpyHostToDevice));
    }
    bd.gas = differs ? b.gas : nullptr;
}

void split_inflow_gas(chkn_sim* s, const std::vector<real>& pTYB)
{
    const size_t ns = pTYB.size() / 3;
    s->inflow_gas.assign(3 * std::max<size_t>(ns, 1), 0.0);
    s->inflow_differs.assign(std::max<size_t>(ns, 1), 0);
    for (size_t q = 0; q < ns; ++q) {
        const real p = pTYB[3 * q], T = pTYB[3 * q + 1], YB = pTYB[3 * q + 2];
        const real* w = s->inflow_w.data() + 6 * q;  // rho, e from update_from_pT
        const real a = eos_a(T);                      // gas.cu:72
        s->inflow_gas[3 * q] = p;
        s->inflow_gas[3 * q + 1] = T;
        s->inflow_gas[3 * q + 2] = a;
        s->inflow_differs[q] = eos_agrees(w[0], w[1], YB, p, T, a) ? 0 : 1;
    }
    dfree(s->d_inflow_gas);
    s->d_inflow_gas = nullptr;
    dalloc(s, (void**)&s->d_inflow_gas, sizeof(real) * s->inflow_gas.size());
    HIPCHK(hipMemcpy(s->d_inflow_gas, s->inflow_gas.data(), sizeof(real) * s->inflow_gas.size(),
                     hipMemcpyHostToDevice));
}

void split_set_inflow(chkn_sim* s, int q)
{
    const real* w = s->inflow_w.data() + 6 * (size_t)q;
    const real T = eos_T(w[1], w[2]);
    const real g[3] = {eos_p(w[0], T), T, eos_a(T)};
    std::copy(g, g + 3, s->inflow_gas.begin() + 3 * (size_t)q);
    s->inflow_differs[q] = 0;
    HIPCHK(hipMemcpy(s->d_inflow_gas + 3 * (size_t)q, g, sizeof g, hipMemcpyHostToDevice));
    for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
}

void split_tables(chkn_sim* s)
{
    dfree(s->d_split_faces);
    dfree(s->d_split_cells);
    dfree(s->d_split_patches);
    s->d_split_faces = nullptr;
    s->d_split_cells = nullptr;
    s->d_split_patches = nullptr;
    s->nsplit_faces = s->nsplit_cells = s->nsplit_patches = 0;
    if (!need_gas(s) && !s->any_reacting) return;
    std::vector<FaceChunk> fc;
    std::vector<CellChunk> cc;
    std::vector<FacePatch> fp;  // viscous pass: block, k, patch, direction
    for (size_t ib = 0; ib < s->blocks.size(); ++ib) {
        const HostBlock& b = s->blocks[ib];
        if (!b.active) continue;
        const long long nf[3] = {(long long)(b.nic + 1) * b.njc * b.nkc,
                                 (long long)b.nic * (b.njc + 1) * b.nkc,
                                 (long long)b.nic * b.njc * (b.nkc + 1)};
        for (int d = 0; d < 3; ++d)
            for (long long f0 = 0; f0 < nf[d]; f0 += kChunk)
                fc.push_back(FaceChunk{(int)ib, d, (int)f0, (int)std::min<long long>(f0 + kChunk, nf[d])});
        for (int c0 = 0; c0 < b.n; c0 += kChunk)
            cc.push_back(CellChunk{(int)ib, c0, std::min(c0 + kChunk, b.n)});
        if (s->viscous)
            for (int k = 0; k <= b.nkc; ++k)
                for (int j0 = 0; j0 <= b.njc; j0 += kPatchJ)
                    for (int i0 = 0; i0 <= b.nic; i0 += kPatchI)
                        for (int d = 0; d < 3; ++d) {
                            if (k == b.nkc && d != 2) continue;          // only k-faces at nkc
                            if (d != 0 && i0 >= b.nic) continue;         // i = nic: i-faces only
                            if (d != 1 && j0 >= b.njc) continue;         // j = njc: j-faces only
                            fp.push_back(FacePatch{(int)ib, d, i0, j0, k});
                        }
    }
    s->nsplit_patches = (int)fp.size();
    if (!fp.empty()) {
        HIPCHK(hipMalloc(&s->d_split_patches, sizeof(FacePatch) * fp.size()));
        HIPCHK(hipMemcpy(s->d_split_patches, fp.data(), sizeof(FacePatch) * fp.size(), hipMemcpyHostToDevice));
    }
    s->nsplit_faces = (int)fc.size();
    s->nsplit_cells = (int)cc.size();
    if (!fc.empty()) {
        HIPCHK(hipMalloc(&s->d_split_faces, sizeof(FaceChunk) * fc.size()));
        HIPCHK(hipMemcpy(s->d_split_faces, fc.data(), sizeof(FaceChunk) * fc.size(), hipMemcpyHostToDevice));
    }
    if (!cc.empty()) {
        HIPCHK(hipMalloc(&s->d_split_cells, sizeof(CellChunk) * cc.size()));
        HIPCHK(hipMemcpy(s->d_split_cells, cc.data(), sizeof(CellChunk) * cc.size(), hipMemcpyHostToDevice));
    }
}

void split_free(chkn_sim* s)
{
    for (auto& b : s->blocks) {
        dfree(b.F); dfree(b.FS); dfree(b.fpos); dfree(b.cpos); dfree(b.gas);
        b.F = b.FS = b.fpos = b.cpos = b.gas = nullptr;
        for (auto& p : b.side) {
            dfree(p);
            p = nullptr;
        }
    }
    dfree(s->d_split_faces); dfree(s->d_split_cells); dfree(s->d_inflow_gas); dfree(s->d_split_patches);
    s->d_split_patches = nullptr;
    s->d_split_faces = nullptr;
    s->d_split_cells = nullptr;
    s->d_inflow_gas = nullptr;
}

long long launches_per_stage(chkn_sim* s, bool file_gas)
{
    if (split_always(s) || file_gas) return s->viscous ? 3 : 2;
    if (!s->opts.per_block_launch) return 1;
    long long n = 0;
    for (auto& b : s->blocks) if (b.active && b.tile_count) n++;
    return n;
}

void launch_stage_form(chkn_sim* s, const StageArgs& a, int stage, bool dtred, bool file_gas,
                       hipStream_t st)
{
    if (split_always(s) || file_gas) {
        SplitArgs sa;
        sa.blocks = a.blocks;
        sa.inflow = a.inflow;
        sa.inflow_gas = s->d_inflow_gas;
        sa.faces = s->d_split_faces;
        sa.patches = s->d_split_patches;
        sa.cells = s->d_split_cells;
        sa.ctl = a.ctl;
        sa.win = a.win;
        sa.wout = a.wout;
        sa.file_gas = file_gas ? 1 : 0;
        HIPCHK(launch_split_faces(sa, s->nsplit_faces, s->x_order, s->flux_calc, s->viscous, st));
        if (s->viscous) HIPCHK(launch_split_viscous(sa, s->nsplit_patches, st));
        HIPCHK(launch_split_update(sa, s->nsplit_cells, stage, dtred, st));
        return;
    }
    StageArgs b = a;
    if (s->opts.per_block_launch) {
        for (auto& hb : s->blocks) {
            if (!hb.active || hb.tile_count == 0) continue;
            b.tile_base = hb.tile_first;
            b.ntiles = hb.tile_count;
            HIPCHK(launch_stage(stage, s->x_order, s->opts.tile_i, s->opts.tile_j, s->opts.vertex_frames,
                                dtred, s->opts.persistent != 0, s->grid_limit, b, st));
        }
    } else {
        b.tile_base = 0;
        b.ntiles = s->ntiles;
        HIPCHK(launch_stage(stage, s->x_order, s->opts.tile_i, s->opts.tile_j, s->opts.vertex_frames,
                            dtred, s->opts.persistent != 0, s->grid_limit, b, st));
    }
}

}  // namespace rt
}  // namespace chkn
