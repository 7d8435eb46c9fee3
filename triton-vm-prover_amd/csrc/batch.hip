// Batched commits of many tables of uneven shapes.
//
// A plan turns a list of table commits (each: LDE -> row digests -> Merkle
// tree, as tvh_commit_dev) into one launch per phase over all tables: the
// three LDE steps, the row hashing and each group of Merkle levels run as
// multi-table kernels whose blocks find their table by binary search over
// per-table block offsets.  The per-phase argument arrays live in one device
// allocation made when the plan is created, so launching a plan allocates
// nothing and records only kernel nodes: it can be captured into a hipGraph
// once and replayed (tvh_batch_plan_launch_graph).
#include <algorithm>
#include <cstring>
#include <vector>

#include "tvh_internal.hpp"

using namespace tvh;

struct PhaseArrays {
    u32 nseg = 0, blocks = 0;
    size_t lds = 0;
    size_t args_off = 0, start_off = 0;  // byte offsets into the plan allocation
};

struct tvh_batch_plan {
    int device = -1;
    void* mem = nullptr;
    PhaseArrays lde[LDE_PHASES];  // indexed by LdePhase, launched in that order
    PhaseArrays hash;
    std::vector<PhaseArrays> merkle;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace {

template <class T>
void append(std::vector<uint8_t>& blob, const std::vector<T>& v, size_t* off) {
    size_t a = (blob.size() + 255) & ~size_t(255);
    blob.resize(a + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(blob.data() + a, v.data(), v.size() * sizeof(T));
    *off = a;
}

template <class T>
void add_phase(std::vector<uint8_t>& blob, PhaseArrays& ph, const std::vector<T>& args,
               const std::vector<u32>& blocks, const std::vector<size_t>& lds) {
    ph.nseg = (u32)args.size();
    std::vector<u32> start(args.size() + 1, 0);
    for (size_t i = 0; i < args.size(); i++) start[i + 1] = start[i] + blocks[i];
    ph.blocks = start.back();
    ph.lds = lds.empty() ? 0 : *std::max_element(lds.begin(), lds.end());
    append(blob, args, &ph.args_off);
    append(blob, start, &ph.start_off);
}

}  // namespace

extern "C" {

int tvh_batch_plan_create(const tvh_table_desc* tables, size_t count, tvh_batch_plan** out) {
    if (!out) return set_error(TVH_ERR_INVALID_ARG, "null plan pointer");
    *out = nullptr;
    if (count && !tables) return set_error(TVH_ERR_INVALID_ARG, "null table list");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    std::vector<Step1Args> s1;
    std::vector<LdeArgs> s2w, s2i, s2e;  // wide row step; split form: inverse rows, expansion
    std::vector<Step3Args> s3;
    std::vector<u32> b1, b2w, b2i, b2e, b3;
    std::vector<size_t> l1, l2w, l2i, l2e, l3;
    std::vector<HashArgs> h;
    std::vector<u32> hb;
    std::vector<std::vector<MerkleArgs>> ma;
    std::vector<std::vector<u32>> mb;
    for (size_t i = 0; i < count; i++) {
        const tvh_table_desc& t = tables[i];
        if (t.layout != TVH_TREE_LEVEL_ORDER && t.layout != TVH_TREE_HEAP)
            return set_error(TVH_ERR_LAYOUT, "table %zu: layout %d", i, t.layout);
        if (!t.codewords || !t.tree || (t.num_cols && !t.trace))
            return set_error(TVH_ERR_INVALID_ARG, "table %zu: null buffer", i);
        if (t.rand_len && !t.rand) return set_error(TVH_ERR_INVALID_ARG, "table %zu: null rand", i);
        if (!is_pow2(t.target_len)) return set_error(TVH_ERR_NOT_POW2, "table %zu: target_len", i);
        RandSpec rs;
        rs.ptr = t.rand_len ? t.rand : nullptr;
        rs.k = rs.ptr ? t.rand_len : 0;
        rs.group_stride = t.rand_len;
        LdePlan P;
        int rc = prepare_lde(t.trace, t.num_cols, t.trace_len, rs, t.trace_offset, t.target_offset,
                             t.target_len, false, t.codewords, &P, true);
        if (rc) return rc;
        if (t.num_cols) {
            if (P.k1) {
                s1.push_back(P.s1);
                b1.push_back(P.s1_blocks);
                l1.push_back(P.s1_lds);
            }
            if (P.split) {
                s2i.push_back(P.s2inv);
                b2i.push_back(P.inv_blocks);
                l2i.push_back(P.inv_lds);
                s2e.push_back(P.s2);
                b2e.push_back(P.s2_blocks);
                l2e.push_back(P.s2_lds);
            } else {
                s2w.push_back(P.s2);
                b2w.push_back(P.s2_blocks);
                l2w.push_back(P.s2_lds);
            }
            if (P.k3) {
                s3.push_back(P.s3);
                b3.push_back(P.s3_blocks);
                l3.push_back(P.s3_lds);
            }
        }
        u64 m = t.target_len;
        u64* leaves = t.tree + 5 * merkle_level_base(m, 0, t.layout);
        h.push_back(HashArgs{t.codewords, m, (u32)t.num_cols, leaves});
        hb.push_back((u32)((m + 255) / 256));
        std::vector<MerkleGroup> groups;
        merkle_groups(t.tree, m, t.layout, groups);
        for (size_t g = 0; g < groups.size(); g++) {
            if (ma.size() <= g) {
                ma.resize(g + 1);
                mb.resize(g + 1);
            }
            ma[g].push_back(groups[g].args);
            mb[g].push_back(groups[g].blocks);
        }
    }
    tvh_batch_plan* plan = new tvh_batch_plan();
    plan->device = ds->device;
    std::vector<uint8_t> blob;
    add_phase(blob, plan->lde[LDE_PHASE_STEP1], s1, b1, l1);
    add_phase(blob, plan->lde[LDE_PHASE_ROWINV], s2i, b2i, l2i);
    add_phase(blob, plan->lde[LDE_PHASE_STEP2], s2w, b2w, l2w);
    add_phase(blob, plan->lde[LDE_PHASE_EXPAND], s2e, b2e, l2e);
    add_phase(blob, plan->lde[LDE_PHASE_STEP3], s3, b3, l3);
    add_phase(blob, plan->hash, h, hb, {});
    plan->merkle.resize(ma.size());
    for (size_t g = 0; g < ma.size(); g++) add_phase(blob, plan->merkle[g], ma[g], mb[g], {});
    if (hipMalloc(&plan->mem, blob.size() ? blob.size() : 256) != hipSuccess) {
        delete plan;
        return set_error(TVH_ERR_OOM, "batch plan arrays");
    }
    if (!blob.empty() &&
        hipMemcpy(plan->mem, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(plan->mem);
        delete plan;
        return set_error(TVH_ERR_DEVICE, "batch plan upload");
    }
    *out = plan;
    return TVH_OK;
}

int tvh_batch_plan_launch(tvh_batch_plan* plan, void* stream) {
    if (!plan) return set_error(TVH_ERR_INVALID_ARG, "null plan");
    hipStream_t st = (hipStream_t)stream;
    auto at = [&](size_t off) { return (const void*)((const uint8_t*)plan->mem + off); };
    int rc = TVH_OK;
    for (int q = 0; q < LDE_PHASES && !rc; q++) {
        const PhaseArrays& ph = plan->lde[q];
        if (ph.nseg)
            rc = launch_lde_multi(q, at(ph.args_off), (const u32*)at(ph.start_off), ph.nseg,
                                  ph.blocks, ph.lds, st);
    }
    if (!rc && plan->hash.nseg)
        rc = launch_hash_rows_multi((const HashArgs*)at(plan->hash.args_off),
                                    (const u32*)at(plan->hash.start_off), plan->hash.nseg,
                                    plan->hash.blocks, st);
    for (size_t g = 0; g < plan->merkle.size() && !rc; g++) {
        const PhaseArrays& ph = plan->merkle[g];
        rc = launch_merkle_multi((const MerkleArgs*)at(ph.args_off), (const u32*)at(ph.start_off),
                                 ph.nseg, ph.blocks, st);
    }
    return rc;
}

int tvh_batch_plan_launch_graph(tvh_batch_plan* plan, void* stream) {
    if (!plan) return set_error(TVH_ERR_INVALID_ARG, "null plan");
    if (!stream) return set_error(TVH_ERR_INVALID_ARG, "graph launch needs a non-null stream");
    hipStream_t st = (hipStream_t)stream;
    if (!plan->exec) {
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "begin capture");
        int rc = tvh_batch_plan_launch(plan, stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess || !g) return set_error(TVH_ERR_DEVICE, "end capture");
        if (hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            plan->exec = nullptr;
            return set_error(TVH_ERR_DEVICE, "graph instantiate");
        }
        plan->graph = g;
    }
    if (hipGraphLaunch(plan->exec, st) != hipSuccess)
        return set_error(TVH_ERR_DEVICE, "graph launch");
    return TVH_OK;
}

int tvh_batch_plan_destroy(tvh_batch_plan* plan) {
    if (!plan) return TVH_OK;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->mem) (void)hipFree(plan->mem);
    delete plan;
    return TVH_OK;
}

int tvh_commit_batch_dev(const tvh_table_desc* tables, size_t count, void* stream) {
    tvh_batch_plan* plan = nullptr;
    int rc = tvh_batch_plan_create(tables, count, &plan);
    if (rc) return rc;
    rc = tvh_batch_plan_launch(plan, stream);
    // the argument arrays must outlive the kernels that read them
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess && !rc)
        rc = set_error(TVH_ERR_DEVICE, "stream sync");
    tvh_batch_plan_destroy(plan);
    return rc;
}

}  // extern "C"
