// Grouped block-sparse permute: one launch moves every subblock of a tensor.
//
// The reference walks sectors x domain trees x codomain trees on the host and, per subblock,
// copies it out of its sector block, shuffles it with Eigen and adds it (scaled by the
// recoupling coefficient) into the output block (src/Core/Tensor.cpp:179-212, with
// PlainInterface::tensor_from_matrix_block / shuffle / add_to_block_from_tensor).
//
// Here the host builds a gather plan once: every output subblock ("task") knows its sources
// and coefficients.  The kernel walks the concatenated output element space in tiles; each
// output element is written exactly once (accumulated over its sources in the reference's
// order without fused multiply-add), so abelian permutes are bit-exact copies/negations.
//
// Tasks whose first output leg maps to a unit-stride source leg are copied with coalesced
// reads and writes.  Tasks where the leading output leg is strided in the source (a real
// transpose) go through a 64x64 LDS tile so that both the global loads and the global stores
// are coalesced.
#include <algorithm>
#include <cstring>

#include "device.hpp"
#include "../common/pool.hpp"

using xh::check;
using xh::guard;

namespace {

constexpr int PT_THREADS = 256;
constexpr int PT_TILE = 2048;  // output elements per workgroup in the gather path
constexpr int TR_DIM = 64;     // transpose tile edge

struct TaskMeta {
    int64_t dst_off, dst_ld;
    uint32_t rows, rmagic, rshift, pad;  // rmagic / rshift: magic-number division by rows
    int32_t nsrc, src_begin;
};
// (row, column) of element el of a task's column-major output subblock, without an integer divide
__device__ __forceinline__ void split_el(const TaskMeta& tm, uint32_t el, uint32_t& row, uint32_t& col)
{
    col = xh::fdiv(el, tm.rmagic, tm.rshift);
    row = el - col * tm.rows;
}

using xh::leg_cstride;
using xh::leg_rstride;
using xh::leg_shift;
using xh::LegInfo;

struct PermDev {
    int nlegs, new_rank;
    int64_t ntasks, total;
    const TaskMeta* meta;
    const LegInfo* legs;  // ntasks * nlegs
    const int64_t* prefix;  // ntasks + 1 (element prefix of gather tasks)
    const int32_t* tile_first;  // gather tiles: first task of each tile (+1 sentinel)
    const xh_permute_source* src;
    // transpose path
    int64_t ntr_tiles;
    const int4* tr_tiles;  // (task, tile index along the contiguous-source leg, tile index along rows, rest index)
};

// Source row / column offsets of the output element (row, col) of a task: every output leg's
// index is peeled off the output row (domain legs) or column (codomain legs) and mapped through
// its source stride.  NLT > 0 fixes the leg count at compile time (the loop unrolls and the
// leg records load together); NLT = 0 reads it from nlegs.
template <int NLT>
__device__ __forceinline__ void decode_src(const LegInfo* __restrict__ L, int nlegs, int new_rank, uint32_t row, uint32_t col, int64_t& srow,
                                           int64_t& scol)
{
    srow = 0;
    scol = 0;
    const int nl = NLT > 0 ? NLT : nlegs;
#pragma unroll
    for(int k = 0; k < (NLT > 0 ? NLT : XH_MAXLEGS); ++k) {
        if(NLT == 0 && k >= nl) break;
        const LegInfo li = L[k];
        const bool is_row = k < new_rank;
        const uint32_t x = is_row ? row : col;
        const uint32_t q = xh::fdiv(x, li.magic, leg_shift(li));
        const uint32_t j = x - q * li.ext;
        if(is_row)
            row = q;
        else
            col = q;
        srow += int64_t(j) * leg_rstride(li);
        scol += int64_t(j) * leg_cstride(li);
    }
}

__device__ __forceinline__ double gather_one(const PermDev& p, const TaskMeta& tm, const double* __restrict__ src, int64_t srow, int64_t scol)
{
    double acc = 0.0;
    for(int s = 0; s < tm.nsrc; ++s) {
        const xh_permute_source so = p.src[tm.src_begin + s];
        const double v = src[so.src_off + srow + scol * so.src_ld];
        acc = __dadd_rn(acc, __dmul_rn(so.coeff, v));
    }
    return acc;
}

// Gather path.  A workgroup covers PT_TILE consecutive output elements (in task order, each task
// column-major in its output subblock); thread i handles elements i, i + PT_THREADS, ...  A tile
// inside one task reads that task's metadata once (uniform across the workgroup); a tile that
// spans several tasks stages their element offsets in LDS and finds each element's task by binary
// search there.
//
// Single-source plans (abelian symmetries) run in three phases: every element's source offset,
// then all PER source loads back to back (independent loads in flight per thread), then the
// scaled stores.  0 + c v keeps the sign of zero of the general path's accumulation.  Plans with
// several sources per task (recoupling sums) accumulate per element in the reference's order.
// NLT: the plan's leg count when it is 2 .. 8, else 0 (read at run time).
constexpr int MT_CAP = 64;  // tasks of a multi-task tile whose records are staged in LDS

template <bool SINGLE_SRC, int NLT>
__global__ __launch_bounds__(PT_THREADS) void k_permute_gather(PermDev p, const double* __restrict__ src, double* __restrict__ dst)
{
    constexpr int PER = PT_TILE / PT_THREADS;
    constexpr int SL = NLT > 0 ? NLT : 1;
    __shared__ int32_t start[PT_TILE + 1];  // tile-relative first element of tasks t_lo .. t_hi (clamped at 0)
    // records of the tile's tasks (fixed leg count, at most MT_CAP tasks): read once per tile
    __shared__ TaskMeta s_meta[MT_CAP];
    __shared__ int64_t s_pre[MT_CAP];
    __shared__ LegInfo s_legs[MT_CAP * SL];
    __shared__ xh_permute_source s_src[SINGLE_SRC ? MT_CAP : 1];
    const int64_t e0 = int64_t(blockIdx.x) * PT_TILE;
    const int n_here = static_cast<int>(min(int64_t(PT_TILE), p.total - e0));
    const int t_lo = p.tile_first[blockIdx.x];
    const int t_hi = p.tile_first[blockIdx.x + 1];
    int64_t soff[PER], dpos[PER];
    double coef[PER];
    double val[PER];
    if(t_lo == t_hi) {
        const TaskMeta tm = p.meta[t_lo];
        const LegInfo* L = p.legs + int64_t(t_lo) * p.nlegs;
        const int64_t base = e0 - p.prefix[t_lo];
        xh_permute_source so{};
        if constexpr(SINGLE_SRC) so = p.src[tm.src_begin];
#pragma unroll
        for(int i = 0; i < PER; ++i) {
            const int r = threadIdx.x + i * PT_THREADS;
            dpos[i] = -1;
            soff[i] = 0;
            coef[i] = 0.0;
            val[i] = 0.0;
            if(r < n_here) {
                const uint32_t el = static_cast<uint32_t>(base + r);
                uint32_t row, col;
                split_el(tm, el, row, col);
                int64_t srow, scol;
                decode_src<NLT>(L, p.nlegs, p.new_rank, row, col, srow, scol);
                dpos[i] = tm.dst_off + row + int64_t(col) * tm.dst_ld;
                if constexpr(SINGLE_SRC) {
                    soff[i] = so.src_off + srow + scol * so.src_ld;
                    coef[i] = so.coeff;
                } else {
                    val[i] = gather_one(p, tm, src, srow, scol);
                }
            }
        }
    } else {
        // tasks t_lo .. t_hi intersect the tile; t_hi - t_lo <= PT_TILE since every task holds an element
        const int nt = t_hi - t_lo + 1;
        for(int j = threadIdx.x; j < nt; j += PT_THREADS) start[j] = static_cast<int32_t>(max(int64_t(0), p.prefix[t_lo + j] - e0));
        const bool staged = NLT > 0 && nt <= MT_CAP;  // uniform over the workgroup
        if(staged) {
            for(int j = threadIdx.x; j < nt; j += PT_THREADS) {
                s_meta[j] = p.meta[t_lo + j];
                s_pre[j] = p.prefix[t_lo + j];
                if constexpr(SINGLE_SRC) s_src[j] = p.src[p.meta[t_lo + j].src_begin];
            }
            for(int j = threadIdx.x; j < nt * SL; j += PT_THREADS) s_legs[j] = p.legs[int64_t(t_lo) * SL + j];
        }
        __syncthreads();
        // the element's task by binary search over the staged offsets; its records come from LDS
        // (staged) or global memory, in two loops so that each reads one address space
        auto find_task = [&](int r) {
            int lo = 0, hi = nt - 1;  // last j with start[j] <= r
            while(lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if(start[mid] <= r)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            return lo;
        };
        if(staged) {
#pragma unroll
            for(int i = 0; i < PER; ++i) {
                const int r = threadIdx.x + i * PT_THREADS;
                dpos[i] = -1;
                soff[i] = 0;
                coef[i] = 0.0;
                val[i] = 0.0;
                if(r < n_here) {
                    const int lo = find_task(r);
                    const TaskMeta tm = s_meta[lo];
                    const uint32_t el = static_cast<uint32_t>(e0 + r - s_pre[lo]);
                    uint32_t row, col;
                split_el(tm, el, row, col);
                    int64_t srow, scol;
                    decode_src<NLT>(s_legs + lo * SL, p.nlegs, p.new_rank, row, col, srow, scol);
                    dpos[i] = tm.dst_off + row + int64_t(col) * tm.dst_ld;
                    if constexpr(SINGLE_SRC) {
                        const xh_permute_source so = s_src[lo];
                        soff[i] = so.src_off + srow + scol * so.src_ld;
                        coef[i] = so.coeff;
                    } else {
                        val[i] = gather_one(p, tm, src, srow, scol);
                    }
                }
            }
        } else {
#pragma unroll
            for(int i = 0; i < PER; ++i) {
                const int r = threadIdx.x + i * PT_THREADS;
                dpos[i] = -1;
                soff[i] = 0;
                coef[i] = 0.0;
                val[i] = 0.0;
                if(r < n_here) {
                    const int t = t_lo + find_task(r);
                    const TaskMeta tm = p.meta[t];
                    const uint32_t el = static_cast<uint32_t>(e0 + r - p.prefix[t]);
                    const LegInfo* L = p.legs + int64_t(t) * p.nlegs;
                    uint32_t row, col;
                split_el(tm, el, row, col);
                    int64_t srow, scol;
                    decode_src<NLT>(L, p.nlegs, p.new_rank, row, col, srow, scol);
                    dpos[i] = tm.dst_off + row + int64_t(col) * tm.dst_ld;
                    if constexpr(SINGLE_SRC) {
                        const xh_permute_source so = p.src[tm.src_begin];
                        soff[i] = so.src_off + srow + scol * so.src_ld;
                        coef[i] = so.coeff;
                    } else {
                        val[i] = gather_one(p, tm, src, srow, scol);
                    }
                }
            }
        }
    }
    if constexpr(SINGLE_SRC) {
        // inactive elements load src[0]: a single-source plan reads at least one source element
#pragma unroll
        for(int i = 0; i < PER; ++i) val[i] = src[soff[i]];
#pragma unroll
        for(int i = 0; i < PER; ++i) val[i] = __dadd_rn(0.0, __dmul_rn(coef[i], val[i]));
    }
#pragma unroll
    for(int i = 0; i < PER; ++i)
        if(dpos[i] >= 0) dst[dpos[i]] = val[i];
}

// Transpose path.  For a task whose output leg 0 is not the unit-stride source leg, let leg u be
// the output leg that reads source stride 1.  A tile covers TR_DIM values of output leg 0 (the
// contiguous direction of the output) times TR_DIM values of leg u (the contiguous direction of
// the source) at one fixed value of every other leg.  Loads run along leg u, stores along leg 0.
struct TrTaskInfo {
    int32_t task, u;  // task id and the output leg with unit source stride
    uint32_t ext0, extu;
    int64_t rest;  // number of combinations of the remaining legs
};

__global__ __launch_bounds__(PT_THREADS) void k_permute_transpose(PermDev p, const TrTaskInfo* __restrict__ info, const double* __restrict__ src,
                                                                  double* __restrict__ dst)
{
    __shared__ double tile[TR_DIM][TR_DIM + 1];
    const int4 tt = p.tr_tiles[blockIdx.x];
    const TrTaskInfo ti = info[tt.x];
    const int t = ti.task;
    const TaskMeta tm = p.meta[t];
    const LegInfo* L = p.legs + int64_t(t) * p.nlegs;
    const int u = ti.u;
    const uint32_t i0 = uint32_t(tt.z) * TR_DIM;  // along output leg 0
    const uint32_t iu = uint32_t(tt.y) * TR_DIM;  // along output leg u
    // Decode the fixed remaining-leg index (every leg except 0 and u) into base offsets.
    uint32_t rest = uint32_t(tt.w);
    int64_t base_srow = 0, base_scol = 0;
    // output position contributions of the remaining legs: row (domain legs) and col (codomain legs)
    int64_t base_row = 0, base_col = 0;
    int64_t rmul = 1, cmul = 1;
    for(int k = 0; k < p.nlegs; ++k) {
        const LegInfo li = L[k];
        const bool is_row = k < p.new_rank;
        int64_t mult = is_row ? rmul : cmul;
        if(k != 0 && k != u) {
            uint32_t q = xh::fdiv(rest, li.magic, leg_shift(li));
            uint32_t j = rest - q * li.ext;
            rest = q;
            base_srow += int64_t(j) * leg_rstride(li);
            base_scol += int64_t(j) * leg_cstride(li);
            if(is_row)
                base_row += int64_t(j) * mult;
            else
                base_col += int64_t(j) * mult;
        }
        if(is_row)
            rmul *= li.ext;
        else
            cmul *= li.ext;
    }
    // output multiplier of leg u
    int64_t umul = 1;
    {
        const bool urow = u < p.new_rank;
        int64_t m = 1;
        for(int k = urow ? 0 : p.new_rank; k < u; ++k) m *= L[k].ext;
        umul = m;
    }
    const bool urow = u < p.new_rank;
    const LegInfo l0 = L[0];
    const LegInfo lu = L[u];
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;  // 0..3
    // load: rows of the tile indexed by leg-0 value a, columns by leg-u value b; b runs along lanes
    for(int a = ty; a < TR_DIM; a += 4) {
        const uint32_t j0 = i0 + a, ju = iu + tx;
        double v = 0.0;
        if(j0 < ti.ext0 && ju < ti.extu) {
            const int64_t srow = base_srow + int64_t(j0) * leg_rstride(l0) + int64_t(ju) * leg_rstride(lu);
            const int64_t scol = base_scol + int64_t(j0) * leg_cstride(l0) + int64_t(ju) * leg_cstride(lu);
            v = gather_one(p, tm, src, srow, scol);
        }
        tile[a][tx] = v;
    }
    __syncthreads();
    // store: a runs along lanes (output leg 0 is the contiguous output direction)
    for(int b = ty; b < TR_DIM; b += 4) {
        const uint32_t j0 = i0 + tx, ju = iu + b;
        if(j0 < ti.ext0 && ju < ti.extu) {
            int64_t row = base_row + j0;  // leg 0 is a domain leg with multiplier 1 (new_rank >= 1 here)
            int64_t col = base_col;
            if(urow)
                row += int64_t(ju) * umul;
            else
                col += int64_t(ju) * umul;
            dst[tm.dst_off + row + col * tm.dst_ld] = tile[tx][b];
        }
    }
}

}  // namespace

struct xh_permute_plan : xh::PlanLife {
    PermDev dev{};
    xh::PackedUpload up;
    int64_t ngather_tiles = 0;
    bool single_src = true;  // every gather task has exactly one source
    double bytes = 0.0;
    const TrTaskInfo* trinfo = nullptr;
};

extern "C" int xh_permute_plan_create(xh_ctx* ctx, const xh_permute_desc* desc, xh_permute_plan** out)
{
    *out = nullptr;
    auto* plan = new xh_permute_plan();
    int rc = guard(ctx, [&] {
        xh::refuse_under_capture(ctx, "plan creation");
        const int NL = desc->nlegs;
        if(NL < 0 || NL > XH_MAXLEGS) throw xh::Error(XH_EINVAL, "permute plan: too many legs");
        if(desc->new_rank < 0 || desc->new_rank > NL) throw xh::Error(XH_EINVAL, "permute plan: invalid new_rank");
        const int64_t NT = desc->ntasks;
        if(NT < 0 || NT > 0x7fffffff || desc->nsources < 0) throw xh::Error(XH_EINVAL, "permute plan: invalid task or source count");
        // pass 1: validate every task and pick its kernel (-1: gather, u > 0: transpose via leg u),
        // in chunks on the host worker pool for large plans; slot and tile0 are first numbered
        // within the chunk and shifted by the chunk's base once all chunks are counted
        std::vector<int32_t> route(static_cast<size_t>(NT));
        std::vector<int64_t> slot(static_cast<size_t>(NT)), tile0(static_cast<size_t>(NT), 0);  // table row; first transpose tile
        struct ChunkCount {
            int64_t ngather = 0, ntr = 0, ntr_tiles = 0, total = 0;
            double moved = 0.0, read = 0.0;
            bool single = true;
        };
        const size_t nchunk = NT >= 16384 ? std::max<size_t>(1, std::min<size_t>(xp::WorkerPool::get().size(), size_t(NT) / 4096)) : 1;
        std::vector<ChunkCount> cc(nchunk);
        auto chunk_lo = [&](size_t c) { return int64_t(size_t(NT) * c / nchunk); };
        auto classify = [&](size_t c) {
            ChunkCount& k = cc[c];
            for(int64_t t = chunk_lo(c); t < chunk_lo(c + 1); ++t) {
                const xh_permute_task& tk = desc->tasks[t];
                if(tk.rows < 0 || tk.cols < 0 || tk.rows * tk.cols >= (int64_t(1) << 31))
                    throw xh::Error(XH_EINVAL, "permute plan: subblock too large for 32-bit indexing");
                if(tk.nsrc < 0 || tk.src_begin < 0 || tk.src_begin + tk.nsrc > desc->nsources)
                    throw xh::Error(XH_EINVAL, "permute plan: source range out of bounds");
                int64_t rprod = 1, cprod = 1;
                for(int q = 0; q < NL; ++q) {
                    if(tk.ext[q] == 0) throw xh::Error(XH_EINVAL, "permute plan: zero extent");
                    if(tk.sstride[q] < 0 || tk.sstride[q] >= (int64_t(1) << 31)) throw xh::Error(XH_EINVAL, "permute plan: source stride out of range");
                    (q < desc->new_rank ? rprod : cprod) *= tk.ext[q];
                }
                if(rprod != tk.rows || cprod != tk.cols) throw xh::Error(XH_EINVAL, "permute plan: leg extents do not match the subblock shape");
                const double elems = double(tk.rows) * double(tk.cols);
                k.moved += elems;
                k.read += elems * tk.nsrc;
                // transpose path when output leg 0 is a real leg (ext >= 16) whose source stride is
                // not 1 and some other leg with ext >= 16 has unit source stride
                int u = -1;
                if(desc->new_rank >= 1 && NL >= 2 && tk.nsrc >= 1 && tk.sstride[0] != 1 && !tk.is_col[0] && tk.ext[0] >= 16) {
                    for(int q = 1; q < NL; ++q)
                        if(tk.sstride[q] == 1 && !tk.is_col[q] && tk.ext[q] >= 16) {
                            u = q;
                            break;
                        }
                }
                route[t] = u;
                if(u <= 0 && tk.nsrc != 1) k.single = false;
                if(u > 0) {
                    int64_t rest = 1;
                    for(int q = 0; q < NL; ++q)
                        if(q != 0 && q != u) rest *= tk.ext[q];
                    slot[t] = k.ntr;  // offset by ngather below
                    tile0[t] = k.ntr_tiles;
                    k.ntr_tiles += rest * ((tk.ext[0] + TR_DIM - 1) / TR_DIM) * ((tk.ext[u] + TR_DIM - 1) / TR_DIM);
                    ++k.ntr;
                } else {
                    slot[t] = k.ngather;
                    ++k.ngather;
                    k.total += tk.rows * tk.cols;
                }
            }
        };
        if(nchunk > 1)
            xp::WorkerPool::get().run(nchunk, classify);
        else
            classify(0);
        std::vector<ChunkCount> base(nchunk);  // per chunk: counts of the chunks before it
        int64_t ngather = 0, ntr = 0, ntr_tiles = 0, total = 0;
        double moved = 0.0, read = 0.0;
        for(size_t c = 0; c < nchunk; ++c) {
            base[c].ngather = ngather;
            base[c].ntr = ntr;
            base[c].ntr_tiles = ntr_tiles;
            base[c].total = total;
            ngather += cc[c].ngather;
            ntr += cc[c].ntr;
            ntr_tiles += cc[c].ntr_tiles;
            total += cc[c].total;
            moved += cc[c].moved;
            read += cc[c].read;
            if(!cc[c].single) plan->single_src = false;
        }
        if(ntr_tiles > 0x7fffffff) throw xh::Error(XH_EINVAL, "permute: too many tiles");
        const int64_t ntiles = (total + PT_TILE - 1) / PT_TILE;
        if(ntiles > 0x7fffffff) throw xh::Error(XH_EINVAL, "permute: too many tiles");
        // pass 2: write the device tables straight into the upload buffer; gather tasks first,
        // then the transpose tasks (their metadata follows the gather metadata)
        const int64_t nmeta = ngather + ntr;
        xh::PackedUpload& up = plan->up;
        const size_t bytes = xh::PackedUpload::packed_bytes(
            {size_t(nmeta) * sizeof(TaskMeta), size_t(nmeta * NL) * sizeof(LegInfo), size_t(ngather + 1) * sizeof(int64_t), size_t(ntiles + 1) * sizeof(int32_t),
             size_t(desc->nsources) * sizeof(xh_permute_source), size_t(ntr_tiles) * sizeof(int4), size_t(ntr) * sizeof(TrTaskInfo)});
        if(bytes >= (size_t(1) << 18)) up.begin_pinned(ctx, bytes);  // large plans: straight into pinned staging
        const size_t o_meta = up.reserve<TaskMeta>(nmeta), o_legs = up.reserve<LegInfo>(nmeta * NL), o_prefix = up.reserve<int64_t>(ngather + 1),
                     o_tiles = up.reserve<int32_t>(ntiles + 1), o_src = up.reserve<xh_permute_source>(desc->nsources),
                     o_tr = up.reserve<int4>(ntr_tiles), o_trinfo = up.reserve<TrTaskInfo>(ntr);
        TaskMeta* meta = up.host_ptr<TaskMeta>(o_meta);
        LegInfo* legs = up.host_ptr<LegInfo>(o_legs);
        int64_t* prefix = up.host_ptr<int64_t>(o_prefix);
        int32_t* tile_first = up.host_ptr<int32_t>(o_tiles);
        int4* trtiles = up.host_ptr<int4>(o_tr);
        TrTaskInfo* trinfo = up.host_ptr<TrTaskInfo>(o_trinfo);
        // second pass over the same chunks: final slots and first transpose tiles, the element
        // prefix of the gather tasks (in table order), the first task of every gather tile (the
        // tiles that start inside a task) and the per-task device tables
        auto place = [&](size_t c) {
            int64_t e = base[c].total;
            for(int64_t t = chunk_lo(c); t < chunk_lo(c + 1); ++t) {
                const xh_permute_task& tk = desc->tasks[t];
                const int u = route[t];
                if(u > 0) {
                    slot[t] += base[c].ntr;
                    tile0[t] += base[c].ntr_tiles;
                } else {
                    slot[t] += base[c].ngather;
                    prefix[slot[t]] = e;
                    const int64_t e_end = e + tk.rows * tk.cols;
                    for(int64_t i = (e + PT_TILE - 1) / PT_TILE; i * PT_TILE < e_end; ++i) tile_first[i] = static_cast<int32_t>(slot[t]);
                    e = e_end;
                }
                const int64_t idx = u > 0 ? ngather + slot[t] : slot[t];
                const xh::FastDiv fd = xh::make_fastdiv(uint32_t(std::max<int64_t>(tk.rows, 1)));
                meta[idx] = TaskMeta{tk.dst_off, tk.dst_ld, uint32_t(tk.rows), fd.magic, fd.shift, 0, tk.nsrc, tk.src_begin};
                LegInfo* L = legs + idx * NL;
                for(int k = 0; k < NL; ++k) L[k] = xh::make_leg(tk.ext[k], tk.sstride[k], tk.is_col[k] != 0);
                if(u > 0) {
                    int64_t rest = 1;
                    for(int k = 0; k < NL; ++k)
                        if(k != 0 && k != u) rest *= tk.ext[k];
                    const int64_t ti = slot[t];
                    trinfo[ti] = TrTaskInfo{static_cast<int32_t>(idx), u, tk.ext[0], tk.ext[u], rest};
                    const int n0 = (tk.ext[0] + TR_DIM - 1) / TR_DIM, nu = (tk.ext[u] + TR_DIM - 1) / TR_DIM;
                    int64_t tt = tile0[t];
                    for(int64_t r = 0; r < rest; ++r)
                        for(int a = 0; a < n0; ++a)
                            for(int b = 0; b < nu; ++b) trtiles[tt++] = int4{static_cast<int>(ti), b, a, static_cast<int>(r)};
                }
            }
        };
        if(nchunk > 1)
            xp::WorkerPool::get().run(nchunk, place);
        else
            place(0);
        prefix[ngather] = total;
        tile_first[ntiles] = static_cast<int32_t>(ngather > 0 ? ngather - 1 : 0);  // sentinel: the last gather task
        if(desc->nsources > 0)
            std::memcpy(up.host_ptr<xh_permute_source>(o_src), desc->sources, size_t(desc->nsources) * sizeof(xh_permute_source));
        plan->dev.nlegs = NL;
        plan->dev.new_rank = desc->new_rank;
        plan->dev.ntasks = nmeta;
        plan->dev.total = total;
        up.upload(ctx);
        plan->dev.meta = up.ptr<const TaskMeta>(o_meta);
        plan->dev.legs = up.ptr<const LegInfo>(o_legs);
        plan->dev.prefix = up.ptr<const int64_t>(o_prefix);
        plan->dev.tile_first = up.ptr<const int32_t>(o_tiles);
        plan->dev.src = up.ptr<const xh_permute_source>(o_src);
        plan->dev.ntr_tiles = ntr_tiles;
        plan->dev.tr_tiles = up.ptr<const int4>(o_tr);
        plan->trinfo = up.ptr<const TrTaskInfo>(o_trinfo);
        plan->ngather_tiles = ntiles;
        plan->bytes = 8.0 * (moved + read);
    });
    if(rc != XH_OK) {
        xh_permute_plan_destroy(ctx, plan);
        return rc;
    }
    *out = plan;
    return XH_OK;
}

namespace {
void permute_plan_release(xh_ctx* ctx, xh::PlanLife* p)
{
    auto* plan = static_cast<xh_permute_plan*>(p);
    plan->up.release(ctx);
    delete plan;
}
}  // namespace

extern "C" int xh_permute_plan_destroy(xh_ctx* ctx, xh_permute_plan* plan)
{
    if(!plan) return XH_OK;
    return guard(ctx, [&] {
        if(plan->pins > 0) {
            plan->dead = true;  // a live graph reads its tables: freed by the last xh_graph_destroy
            return;
        }
        permute_plan_release(ctx, plan);
    });
}

extern "C" int xh_permute_plan_bytes(const xh_permute_plan* plan, double* bytes)
{
    if(!plan || !bytes) return XH_EINVAL;
    *bytes = plan->bytes;
    return XH_OK;
}

namespace {
template <bool SINGLE_SRC>
void launch_gather_s(xh_ctx* ctx, const xh_permute_plan* plan, const double* src, double* dst)
{
    const dim3 grid(static_cast<unsigned>(plan->ngather_tiles)), block(PT_THREADS);
    switch(plan->dev.nlegs) {
    case 2: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 2>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    case 3: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 3>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    case 4: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 4>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    case 5: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 5>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    case 6: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 6>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    case 7: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 7>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    case 8: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 8>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    default: hipLaunchKernelGGL((k_permute_gather<SINGLE_SRC, 0>), grid, block, 0, ctx->stream, plan->dev, src, dst); break;
    }
}

void launch_gather(xh_ctx* ctx, const xh_permute_plan* plan, const double* src, double* dst)
{
    if(plan->single_src)
        launch_gather_s<true>(ctx, plan, src, dst);
    else
        launch_gather_s<false>(ctx, plan, src, dst);
}
}  // namespace

extern "C" int xh_permute(xh_ctx* ctx, const xh_permute_plan* plan, const double* src, double* dst)
{
    return guard(ctx, [&] {
        if(!plan) throw xh::Error(XH_EINVAL, "null permute plan");
        if(plan->dead) throw xh::Error(XH_EINVAL, "permute plan used after xh_permute_plan_destroy");
        xh::pin_for_capture(ctx, const_cast<xh_permute_plan*>(plan), permute_plan_release);
        const bool timed = ctx->perm_timing && (plan->ngather_tiles > 0 || plan->dev.ntr_tiles > 0);
        if(timed) xh::perm_timer_begin(ctx);
        if(plan->ngather_tiles > 0) {
            if(plan->ngather_tiles > 0x7fffffff) throw xh::Error(XH_EINVAL, "permute: too many tiles");
            launch_gather(ctx, plan, src, dst);
            xh::launch_check("k_permute_gather");
        }
        if(plan->dev.ntr_tiles > 0) {
            if(plan->dev.ntr_tiles > 0x7fffffff) throw xh::Error(XH_EINVAL, "permute: too many tiles");
            hipLaunchKernelGGL(k_permute_transpose, dim3(static_cast<unsigned>(plan->dev.ntr_tiles)), dim3(PT_THREADS), 0, ctx->stream, plan->dev,
                               plan->trinfo, src, dst);
            xh::launch_check("k_permute_transpose");
        }
        if(timed) xh::perm_timer_end(ctx, plan->bytes);
    });
}
