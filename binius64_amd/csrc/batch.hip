// Commit jobs: one FRI commit phase (encode, Merkle commit, every fold round and
// its commitment) with pre-seeded challenges, run either alone (plain per-job
// launches) or many at once (batched: each launch covers the same stage of every
// job, and the whole launch sequence is captured in a hipGraph).
//
// A job of 2^n message elements uses the Verifier::setup parameters
// (verifier/verifier/src/verify.rs:61-106): log_inv_rate 1, batch 2^4, arity 4.
// Its stages follow FRIFoldProver::execute_fold_round (prover fri/fold.rs:117-175):
//   NTT passes -> 16-element leaves + interleaved fold -> tree 0
//   -> for each folded oracle k: leaves (coset 2^arity_k or 2^n_final) -> tree
//   -> fold_codeword with the next arity's challenges -> next oracle.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "abi_util.hpp"
#include "host_util.hpp"
#include "internal.hpp"

using b64::abi::as_stream;
using b64::abi::to_u4;
using b64::host::u128;

namespace {

constexpr size_t ALIGN = 256;
inline size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }


b64::CompressIV job_iv() {
    static b64::CompressIV iv;
    static std::once_flag once;
    std::call_once(once, [] {
        static const char tag[] = "BINIUS SHA-256 COMPRESS";
        uint8_t d[32];
        b64::host::Sha256::digest((const uint8_t *)tag, sizeof(tag) - 1, d);
        for (int k = 0; k < 8; k++)
            iv.w[k] = (uint32_t)d[4 * k] | ((uint32_t)d[4 * k + 1] << 8) | ((uint32_t)d[4 * k + 2] << 16) |
                      ((uint32_t)d[4 * k + 3] << 24);
    });
    return iv;
}

// Parameters and prover domain per message size, shared by every job of that size.
struct SizeInfo {
    b64_fri_params p;
    b64_domain *dom = nullptr;
    b64::DeviceDomain dd = {0, nullptr};
};

std::mutex g_size_mu;
std::map<uint32_t, SizeInfo *> g_sizes;

int size_info(uint32_t n, SizeInfo **out) {
    if (n < 8 || n > 40) return B64_ERR_INVALID_ARGS;  // batch 2^4 and at least 4 NTT layers
    std::lock_guard<std::mutex> g(g_size_mu);
    auto it = g_sizes.find(n);
    if (it != g_sizes.end()) {
        *out = it->second;
        return B64_OK;
    }
    SizeInfo *si = new (std::nothrow) SizeInfo();
    if (!si) return B64_ERR_OUT_OF_MEMORY;
    int rc = b64_setup_commit_params(n, 1, 96, &si->p, &si->dom);
    if (rc == B64_OK && si->p.log_batch_size != 4) rc = B64_ERR_INVALID_ARGS;
    if (rc != B64_OK) {
        if (si->dom) b64_domain_destroy(si->dom);
        delete si;
        return rc;
    }
    g_sizes[n] = si;
    *out = si;
    return B64_OK;
}

// The size's twiddle basis on the device (uploaded on first use).
int size_device(SizeInfo *si) {
    if (si->dd.W) return B64_OK;
    return b64_domain_device_view_internal(si->dom, &si->dd);
}

// Device workspace layout of one job.
struct Layout {
    const SizeInfo *si;
    int rs, lbs, n_oracles;  // folded oracles = n_fold_arities + 1
    size_t cw, tree0;
    size_t oracle[B64_MAX_FOLD_ARITIES + 1], tree[B64_MAX_FOLD_ARITIES + 1];
    int olen[B64_MAX_FOLD_ARITIES + 1], ocoset[B64_MAX_FOLD_ARITIES + 1];
    size_t total;
};

void layout_of(const SizeInfo *si, Layout &L) {
    const b64_fri_params &p = si->p;
    L.si = si;
    L.rs = (int)(p.log_dim + p.log_inv_rate);
    L.lbs = (int)p.log_batch_size;
    L.n_oracles = (int)p.n_fold_arities + 1;
    size_t at = 0;
    L.cw = at;
    at = align_up(at + (16ull << (L.rs + L.lbs)));
    L.tree0 = at;
    at = align_up(at + 32ull * ((2ull << L.rs) - 1));
    int len = L.rs;
    for (int k = 0; k < L.n_oracles; k++) {
        if (k > 0) len -= (int)p.fold_arities[k - 1];
        const int coset = k < (int)p.n_fold_arities ? (int)p.fold_arities[k] : (int)p.log_terminal_dim;
        L.olen[k] = len;
        L.ocoset[k] = coset;
        L.oracle[k] = at;
        at = align_up(at + (16ull << len));
        L.tree[k] = at;
        at = align_up(at + 32ull * ((2ull << (len - coset)) - 1));
    }
    L.total = at;
}

// A stage of a job, tagged with its slot in the common launch order.
struct Stage {
    int slot;
    b64::MultiKind kind;
    union {
        b64::NttPassArgs ntt;
        b64::NttJobArgs nttj;
        b64::Leaf16Args leaf16;
        b64::LeafArgs leaf;
        b64::SubtreeArgs sub;
        b64::FoldCwArgs fold;
        b64::FoldIlArgs foldil;
    };
    uint32_t blocks;
    Stage() {}
};

// Slot numbering: 0..15 NTT passes, 16 leaves+fold, 17..30 tree-0 steps, then
// per folded oracle k (base 32 + 16k): 0 fold_codeword, 1 16-element leaves,
// 2 other leaves, 4.. tree steps.  Few long leaves (one block each: only a
// terminal oracle has them, the last stage of its job) and their tree steps go
// after every oracle's slots, so the jobs' terminal oracles, whatever their k,
// share one launch: each is a chain of ~33 compressions per leaf that leaves the
// device nearly idle for ~0.1 ms.
constexpr int SLOT_LEAF16 = 16, SLOT_TREE0 = 17, SLOT_ORACLE = 32;
constexpr int SLOT_TERMINAL = SLOT_ORACLE + 16 * (B64_MAX_FOLD_ARITIES + 1);

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

// The forward NTT passes of an encode (the first gathering the bit-reversed message).
int ntt_stages(const Layout &L, const void *d_msg, uint32_t m, uint4 *cw, std::vector<Stage> &out) {
    const SizeInfo *si = L.si;
    const int log_d = L.rs + L.lbs;
    int starts[64], ks[64], cs[64], es[64];
    const b64::NttTuning &tu = b64::batch_tuning_for(log_d);
    const int np = b64::ntt_pass_plan(tu, log_d, 1, L.lbs, true, starts, ks, cs, es);
    if (np > 16) return B64_ERR_INVALID_ARGS;
    for (int q = 0; q < np; q++) {
        Stage st;
        st.slot = q;
        b64::NttPassArgs a;
        a.data = cw;
        a.W = si->dd.W;
        a.X = si->dd.X;
        a.x_layers = si->dd.x_layers;
        a.msg = q == 0 ? (const uint4 *)d_msg : nullptr;
        a.m = (int)m;
        a.log_d = log_d;
        a.layer0 = starts[q];
        a.k = ks[q];
        a.c = cs[q];
        a.e = es[q];
        uint64_t per_block = 0;
        b64::ntt_multi_pass(tu, log_d, starts[q], ks[q], cs[q], es[q], q == 0, &st.kind, &per_block, &st.blocks);
        if (st.kind == b64::MultiKind::NttPass) {
            // the multi-job table kernel (32-row tables); variant 0 maps to "no
            // layer qualifies for a table", variants >= 3 keep its own threshold
            a.tab_min = tu.variant == 0 ? (1 << 30) : tu.variant >= 3 ? 128 : tu.tab_min;
            st.ntt = a;
        } else {
            a.tab_min = tu.tab_min;
            st.nttj.a = a;
            st.nttj.per_block = per_block;
        }
        out.push_back(st);
    }
    return B64_OK;
}

// 16-element leaves of the codeword (with the interleaved fold into oracle 0
// when chal is given) at SLOT_LEAF16.
void leaf16_stage(const Layout &L, uint8_t *ws, const uint64_t *chal, std::vector<Stage> &out) {
    Stage st;
    st.slot = SLOT_LEAF16;
    st.kind = b64::MultiKind::Leaf16Fold;
    st.leaf16.cw = (const uint4 *)(ws + L.cw);
    st.leaf16.n_leaves = 1ull << L.rs;
    st.leaf16.digests = (uint4 *)(ws + L.tree0);
    st.leaf16.folded = chal ? (uint4 *)(ws + L.oracle[0]) : nullptr;
    for (int t = 0; t < 4; t++) st.leaf16.chal[t] = chal ? to_u4(chal + 2 * t) : make_uint4(0, 0, 0, 0);
    st.blocks = blocks_for(st.leaf16.n_leaves);
    out.push_back(st);
}

// The tree levels above 2^k leaves' digests, from slot0 on; its root is listed.
// Fails when the steps would run past max_slot (the next stage's slot).
int tree_stages(uint4 *tree, uint64_t n_leaves, int slot0, int max_slot, std::vector<Stage> &out,
                std::vector<const uint4 *> &roots) {
    uint64_t offs[32], ns[32];
    int lv[32], qd[32];
    if (slot0 + b64::merkle_reduce_steps(n_leaves, nullptr, nullptr, nullptr, nullptr) > max_slot)
        return B64_ERR_INVALID_ARGS;
    const int steps = b64::merkle_reduce_steps(n_leaves, offs, ns, lv, qd);
    for (int i = 0; i < steps; i++) {
        Stage st;
        st.slot = slot0 + i;
        st.kind = qd[i] ? b64::MultiKind::MerkleQuad : b64::MultiKind::Subtree;
        st.sub.tree = tree;
        st.sub.in_off = offs[i];
        st.sub.n_in = ns[i];
        st.sub.levels = lv[i];
        st.sub.iv = job_iv();
        st.blocks = b64::merkle_step_blocks(ns[i], qd[i]);
        out.push_back(st);
    }
    roots.push_back(tree + 2 * (2 * n_leaves - 2));
    return B64_OK;
}

// Folded oracle k: for k > 0 fold_codeword of oracle k-1 with chal (the
// arity's challenges), then its leaves (cosets) and tree.
int oracle_stages(const Layout &L, uint8_t *ws, int k, const uint64_t *chal, std::vector<Stage> &out,
                  std::vector<const uint4 *> &roots) {
    const b64_fri_params &p = L.si->p;
    const int base = SLOT_ORACLE + 16 * k;
    uint4 *ok = (uint4 *)(ws + L.oracle[k]);
    if (k > 0) {
        const int a = (int)p.fold_arities[k - 1];
        if (a > 4) return B64_ERR_INVALID_ARGS;
        Stage st;
        st.slot = base;
        st.kind = b64::MultiKind::FoldCodeword;
        st.fold.cw = (const uint4 *)(ws + L.oracle[k - 1]);
        st.fold.out = ok;
        st.fold.n_out = 1ull << L.olen[k];
        st.fold.log_len = L.olen[k - 1];
        st.fold.n_chal = a;
        st.fold.W = L.si->dd.W;
        for (int t = 0; t < a; t++) st.fold.chal[t] = to_u4(chal + 2 * t);
        st.blocks = blocks_for(st.fold.n_out);
        out.push_back(st);
    }
    uint4 *tk = (uint4 *)(ws + L.tree[k]);
    const uint64_t n_leaves = 1ull << (L.olen[k] - L.ocoset[k]);
    Stage st;
    st.slot = L.ocoset[k] == 4 ? base + 1 : base + 2;  // one kernel kind per slot
    if (L.ocoset[k] == 4) {
        st.kind = b64::MultiKind::Leaf16Fold;
        st.leaf16.cw = ok;
        st.leaf16.n_leaves = n_leaves;
        st.leaf16.digests = tk;
        st.leaf16.folded = nullptr;
        for (int t = 0; t < 4; t++) st.leaf16.chal[t] = make_uint4(0, 0, 0, 0);
    } else {
        st.kind = b64::MultiKind::LeafHash;
        st.leaf.elems = ok;
        st.leaf.n_leaves = n_leaves;
        st.leaf.batch = 1ull << L.ocoset[k];
        st.leaf.digests = tk;
        if (b64::leaf_hash_long_fits(st.leaf.batch, n_leaves)) {
            st.kind = b64::MultiKind::LeafHashLong;
            st.slot = SLOT_TERMINAL;  // one kernel kind per slot
        }
    }
    const bool terminal = st.kind == b64::MultiKind::LeafHashLong;
    st.blocks = terminal ? (uint32_t)n_leaves : blocks_for(n_leaves);
    out.push_back(st);
    const int tree_base = terminal ? SLOT_TERMINAL - 3 : base;  // tree steps from tree_base + 4
    B64_TRY(tree_stages(tk, n_leaves, tree_base + 4, tree_base + 16, out, roots));
    return B64_OK;
}

int job_stages(const b64_job_desc &jd, const Layout &L, std::vector<Stage> &out, std::vector<const uint4 *> &roots) {
    const b64_fri_params &p = L.si->p;
    uint8_t *ws = (uint8_t *)jd.d_workspace;
    B64_TRY(ntt_stages(L, jd.d_msg, jd.log_msg_len, (uint4 *)(ws + L.cw), out));
    // leaves of the codeword + interleaved fold into oracle 0, then tree 0
    leaf16_stage(L, ws, jd.challenges, out);
    B64_TRY(tree_stages((uint4 *)(ws + L.tree0), 1ull << L.rs, SLOT_TREE0, SLOT_ORACLE - 1, out, roots));
    // folded oracles: fold (k > 0), leaves, tree
    int chal_at = L.lbs;
    for (int k = 0; k < L.n_oracles; k++) {
        B64_TRY(oracle_stages(L, ws, k, jd.challenges + 2 * chal_at, out, roots));
        if (k > 0) chal_at += (int)p.fold_arities[k - 1];
    }
    return B64_OK;
}

// Tables up to this size go up through a copy kernel reading coherent pinned
// host memory: in a kernel trace, a DMA copy ahead of a phase's first kernel
// showed ~30 us between the copy's end and the kernel's start.
constexpr size_t KERNEL_UPLOAD_MAX = 1 << 20;

__global__ __launch_bounds__(256) void k_upload16(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                  uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Tables of `bytes` from host (coherent pinned, at least bytes rounded up to 16)
// to dev (as large), in order on s.
int upload_tables(const uint8_t *host, uint8_t *dev, size_t bytes, hipStream_t s) {
    const size_t up = (bytes + 15) & ~(size_t)15;
    if (up <= KERNEL_UPLOAD_MAX) {
        const uint32_t n = (uint32_t)(up / 16);
        if (n) k_upload16<<<(n + 255) / 256, 256, 0, s>>>((const uint4 *)host, (uint4 *)dev, n);
        HIP_TRY(hipGetLastError());
        return B64_OK;
    }
    HIP_TRY(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
    return B64_OK;
}

// The roots of a plan's jobs to dst (device memory, or coherent pinned host memory).
__global__ void k_gather_roots(const uint4 *const *__restrict__ src, uint4 *__restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        dst[2 * i] = src[i][0];
        dst[2 * i + 1] = src[i][1];
    }
}

}  // namespace

namespace {

// One multi-job launch: a stage slot over every job that has it.
struct Launch {
    b64::MultiKind kind;
    size_t args_off, start_off;  // into the device tables
    int n_jobs;
    uint32_t blocks;
    int fold_nc;  // FoldCodeword: the n_chal all jobs share (1..4), else 0
};

typedef std::map<std::pair<int, int>, std::vector<Stage>> SlotMap;

size_t put_bytes(std::vector<uint8_t> &host, const void *p, size_t n) {
    const size_t off = align_up(host.size());
    host.resize(off + n);
    memcpy(host.data() + off, p, n);
    return off;
}

// Per slot (in slot order) an args array and a block-start array, appended to host.
int build_launches(const SlotMap &slots, std::vector<uint8_t> &host, std::vector<Launch> &launches) {
    for (auto &kv : slots) {
        const std::vector<Stage> &v = kv.second;
        // stages in one slot share a kind
        Launch l;
        l.kind = v[0].kind;
        l.n_jobs = (int)v.size();
        l.fold_nc = l.kind == b64::MultiKind::FoldCodeword ? v[0].fold.n_chal : 0;
        std::vector<uint32_t> start;
        uint32_t total = 0;
        std::vector<uint8_t> args;
        for (const Stage &x : v) {
            if (x.kind != l.kind) return B64_ERR_INVALID_ARGS;
            if (x.kind == b64::MultiKind::FoldCodeword && x.fold.n_chal != l.fold_nc) l.fold_nc = 0;
            start.push_back(total);
            total += x.blocks;
            const void *a = nullptr;
            size_t sz = 0;
            switch (x.kind) {
                case b64::MultiKind::NttPass: a = &x.ntt; sz = sizeof(x.ntt); break;
                case b64::MultiKind::NttR16:
                case b64::MultiKind::NttR16Msg:
                case b64::MultiKind::NttR16W8:
                case b64::MultiKind::NttR16W8Msg:
                case b64::MultiKind::NttT4U9:
                case b64::MultiKind::NttT4U10: a = &x.nttj; sz = sizeof(x.nttj); break;
                case b64::MultiKind::Leaf16Fold: a = &x.leaf16; sz = sizeof(x.leaf16); break;
                case b64::MultiKind::LeafHash:
                case b64::MultiKind::LeafHashLong: a = &x.leaf; sz = sizeof(x.leaf); break;
                case b64::MultiKind::Subtree:
                case b64::MultiKind::MerkleQuad: a = &x.sub; sz = sizeof(x.sub); break;
                case b64::MultiKind::FoldCodeword: a = &x.fold; sz = sizeof(x.fold); break;
                case b64::MultiKind::FoldInterleaved: a = &x.foldil; sz = sizeof(x.foldil); break;
            }
            const size_t o = args.size();
            args.resize(o + sz);
            memcpy(args.data() + o, a, sz);
        }
        l.blocks = total;
        l.args_off = put_bytes(host, args.data(), args.size());
        l.start_off = put_bytes(host, start.data(), start.size() * sizeof(uint32_t));
        launches.push_back(l);
    }
    return B64_OK;
}

int enqueue_launches(const std::vector<Launch> &launches, const uint8_t *d_tables, hipStream_t s) {
    for (const Launch &l : launches)
        HIP_TRY(b64::launch_multi(l.kind, d_tables + l.args_off, (const uint32_t *)(d_tables + l.start_off), l.n_jobs,
                                  l.blocks, s, l.fold_nc));
    return B64_OK;
}

}  // namespace

struct b64_batch_plan {
    std::vector<Launch> launches;
    std::vector<b64_job_desc> jobs;
    std::vector<int> roots_per_job;
    int n_roots = 0;
    uint8_t *d_tables = nullptr;
    const uint4 **d_root_src = nullptr;
    uint8_t *h_roots = nullptr;  // coherent pinned host memory: the roots are written there by a kernel
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = false;
    bool owns_buffers = true;
    // A plan of one job without a graph runs the job's single-job launchers
    // (job_enqueue): with nothing to batch, the per-size schedule is the
    // faster one (a served job at low load is a batch of one).
    bool single = false;

    ~b64_batch_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (owns_buffers) {
            if (d_tables) (void)hipFree(d_tables);
            if (h_roots) (void)hipHostFree(h_roots);
        }
    }

    int enqueue(hipStream_t s) const {
        B64_TRY(enqueue_launches(launches, d_tables, s));
        if (n_roots) {
            k_gather_roots<<<(n_roots + 63) / 64, 64, 0, s>>>(d_root_src, (uint4 *)h_roots, n_roots);
            HIP_TRY(hipGetLastError());
        }
        return B64_OK;
    }
};

extern "C" {

size_t b64_job_workspace_size(uint32_t log_msg_len) {
    SizeInfo *si = nullptr;
    if (log_msg_len < 8 || size_info(log_msg_len, &si) != B64_OK) return 0;
    Layout L;
    layout_of(si, L);
    return L.total;
}

uint32_t b64_job_n_roots(uint32_t log_msg_len) {
    SizeInfo *si = nullptr;
    if (log_msg_len < 8 || size_info(log_msg_len, &si) != B64_OK) return 0;
    return si->p.n_fold_arities + 2;
}

int b64_job_oracle(const b64_job_desc *job, uint32_t k, const void **d_codeword, uint32_t *log_len) {
    if (!job) return B64_ERR_INVALID_ARGS;
    SizeInfo *si = nullptr;
    B64_TRY(size_info(job->log_msg_len, &si));
    Layout L;
    layout_of(si, L);
    if (k > (uint32_t)L.n_oracles) return B64_ERR_INDEX_OUT_OF_RANGE;
    const uint8_t *ws = (const uint8_t *)job->d_workspace;
    if (k == 0) {
        if (d_codeword) *d_codeword = ws + L.cw;
        if (log_len) *log_len = (uint32_t)(L.rs + L.lbs);
    } else {
        if (d_codeword) *d_codeword = ws + L.oracle[k - 1];
        if (log_len) *log_len = (uint32_t)L.olen[k - 1];
    }
    return B64_OK;
}

}  // extern "C"

// A job's root addresses, passed by value to the kernel that writes them out.
struct RootPtrs {
    const uint4 *p[B64_MAX_FOLD_ARITIES + 2];
};

__global__ void k_roots_out(RootPtrs src, int n, uint4 *__restrict__ dst) {
    const int i = threadIdx.x;
    if (i < n) {
        dst[2 * i] = src.p[i][0];
        dst[2 * i + 1] = src.p[i][1];
    }
}

// Plain path: the job's stages as single-job launches on `s` (the per-size
// NTT schedule, register-resident passes included), then its roots into
// roots_dst: written by one kernel when roots_dst is coherent pinned host
// memory (direct), else copied one by one (host memory; pinned for the copies
// to stay asynchronous).  Nothing waits.
static int job_enqueue(const b64_job_desc *job, hipStream_t s, uint8_t *roots_dst, bool direct = false) {
    if (!job || !job->d_msg || !job->d_workspace || !job->challenges || !roots_dst) return B64_ERR_INVALID_ARGS;
    SizeInfo *si = nullptr;
    B64_TRY(size_info(job->log_msg_len, &si));
    B64_TRY(size_device(si));
    Layout L;
    layout_of(si, L);
    std::vector<Stage> stages;
    std::vector<const uint4 *> roots;
    B64_TRY(job_stages(*job, L, stages, roots));
    const b64_fri_params &p = si->p;
    uint8_t *ws = (uint8_t *)job->d_workspace;
    // the same computation through the single-job launchers
    HIP_TRY(b64::launch_ntt(si->dd, (uint4 *)(ws + L.cw), L.rs + L.lbs, 1, L.lbs, false, s,
                            (const uint4 *)job->d_msg, (int)job->log_msg_len));
    uint4 ch[4];
    for (int t = 0; t < 4; t++) ch[t] = to_u4(job->challenges + 2 * t);
    HIP_TRY(b64::launch_leaf16_fold((const uint4 *)(ws + L.cw), 1ull << L.rs, (uint4 *)(ws + L.tree0),
                                    (uint4 *)(ws + L.oracle[0]), ch, s));
    HIP_TRY(b64::launch_merkle_levels((uint4 *)(ws + L.tree0), 1ull << L.rs, job_iv(), s));
    int chal_at = L.lbs;
    for (int k = 0; k < L.n_oracles; k++) {
        if (k > 0) {
            const int a = (int)p.fold_arities[k - 1];
            uint4 c[b64::MAX_FOLD_CHAL];
            for (int t = 0; t < a; t++) c[t] = to_u4(job->challenges + 2 * (chal_at + t));
            chal_at += a;
            HIP_TRY(b64::launch_fold_codeword(si->dd, (const uint4 *)(ws + L.oracle[k - 1]), (uint4 *)(ws + L.oracle[k]),
                                              L.olen[k - 1], c, a, s));
        }
        const uint64_t n_leaves = 1ull << (L.olen[k] - L.ocoset[k]);
        HIP_TRY(b64::launch_leaf_hash((const uint4 *)(ws + L.oracle[k]), n_leaves, 0, 1ull << L.ocoset[k],
                                      (uint4 *)(ws + L.tree[k]), s));
        HIP_TRY(b64::launch_merkle_levels((uint4 *)(ws + L.tree[k]), n_leaves, job_iv(), s));
    }
    if (direct && roots.size() <= B64_MAX_FOLD_ARITIES + 2) {
        RootPtrs rp;
        for (size_t r = 0; r < roots.size(); r++) rp.p[r] = roots[r];
        k_roots_out<<<1, 64, 0, s>>>(rp, (int)roots.size(), (uint4 *)roots_dst);
        HIP_TRY(hipGetLastError());
        return B64_OK;
    }
    for (size_t r = 0; r < roots.size(); r++)
        HIP_TRY(hipMemcpyAsync(roots_dst + 32 * r, roots[r], 32, hipMemcpyDeviceToHost, s));
    return B64_OK;
}

extern "C" {

int b64_job_run(const b64_job_desc *job, void *stream) {
    if (!job || !job->roots_out) return B64_ERR_INVALID_ARGS;
    const hipStream_t s = as_stream(stream);
    B64_TRY(job_enqueue(job, s, job->roots_out));
    HIP_TRY(hipStreamSynchronize(s));
    return B64_OK;
}

int b64_batch_plan_create(const b64_job_desc *jobs, uint32_t n_jobs, int use_graph, void *stream,
                          b64_batch_plan **out) {
    return b64_batch_plan_build(jobs, n_jobs, use_graph, nullptr, as_stream(stream), out);
}

}  // extern "C"

int b64_batch_plan_build(const b64_job_desc *jobs, uint32_t n_jobs, int use_graph, b64::BatchStorage *st,
                         hipStream_t upload_stream, b64_batch_plan **out) {
    if (!jobs || !n_jobs || !out) return B64_ERR_INVALID_ARGS;
    b64_batch_plan *plan = new (std::nothrow) b64_batch_plan();
    if (!plan) return B64_ERR_OUT_OF_MEMORY;
    std::unique_ptr<b64_batch_plan> guard(plan);
    plan->jobs.assign(jobs, jobs + n_jobs);
    if (n_jobs == 1 && !use_graph) {
        const b64_job_desc &jd = jobs[0];
        if (!jd.d_msg || !jd.d_workspace || !jd.challenges || !jd.roots_out) return B64_ERR_INVALID_ARGS;
        SizeInfo *si = nullptr;
        B64_TRY(size_info(jd.log_msg_len, &si));
        B64_TRY(size_device(si));
        Layout L;
        layout_of(si, L);
        std::vector<Stage> st_list;
        std::vector<const uint4 *> roots;
        B64_TRY(job_stages(jd, L, st_list, roots));
        plan->single = true;
        plan->n_roots = (int)roots.size();
        plan->roots_per_job.push_back(plan->n_roots);
        if (st) {
            if ((size_t)plan->n_roots > st->roots_cap) return B64_ERR_BUFFER_TOO_SMALL;
            plan->owns_buffers = false;
            plan->h_roots = st->h_roots;
        } else {
            HIP_TRY(hipHostMalloc((void **)&plan->h_roots, 32ull * plan->n_roots, hipHostMallocCoherent));
        }
        *out = guard.release();
        return B64_OK;
    }
    // per-slot job lists; an NTT slot splits by kernel (jobs of different sizes
    // may run a pass in different kernels)
    SlotMap slots;
    std::vector<const uint4 *> root_src;
    for (uint32_t j = 0; j < n_jobs; j++) {
        const b64_job_desc &jd = jobs[j];
        if (!jd.d_msg || !jd.d_workspace || !jd.challenges || !jd.roots_out) return B64_ERR_INVALID_ARGS;
        SizeInfo *si = nullptr;
        B64_TRY(size_info(jd.log_msg_len, &si));
        B64_TRY(size_device(si));
        Layout L;
        layout_of(si, L);
        std::vector<Stage> st;
        std::vector<const uint4 *> roots;
        B64_TRY(job_stages(jd, L, st, roots));
        for (const Stage &x : st) slots[{x.slot, (int)x.kind}].push_back(x);
        plan->roots_per_job.push_back((int)roots.size());
        root_src.insert(root_src.end(), roots.begin(), roots.end());
    }
    plan->n_roots = (int)root_src.size();
    // device tables: per launch an args array and a block-start array
    std::vector<uint8_t> host;
    B64_TRY(build_launches(slots, host, plan->launches));
    const size_t root_src_off = put_bytes(host, root_src.data(), root_src.size() * sizeof(void *));
    if (st) {
        // caller-owned buffers, uploaded in stream order (no allocation, no device sync)
        if (((host.size() + 15) & ~(size_t)15) > st->tables_cap || (size_t)plan->n_roots > st->roots_cap)
            return B64_ERR_BUFFER_TOO_SMALL;
        plan->owns_buffers = false;
        plan->d_tables = st->d_tables;
        plan->h_roots = st->h_roots;
        memcpy(st->h_stage, host.data(), host.size());
        B64_TRY(upload_tables(st->h_stage, plan->d_tables, host.size(), upload_stream));
    } else {
        HIP_TRY(hipMalloc(&plan->d_tables, host.size() ? host.size() : 16));
        HIP_TRY(hipMemcpy(plan->d_tables, host.data(), host.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipHostMalloc((void **)&plan->h_roots, 32ull * (plan->n_roots ? plan->n_roots : 1),
                              hipHostMallocCoherent));
    }
    plan->d_root_src = (const uint4 **)(plan->d_tables + root_src_off);
    if (use_graph) {
        hipStream_t cap = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        int rc = B64_OK;
        if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = B64_ERR_HIP;
        if (rc == B64_OK) {
            rc = plan->enqueue(cap);
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(cap, &g) != hipSuccess) rc = B64_ERR_HIP;
            plan->graph = g;
        }
        if (rc == B64_OK && hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0) != hipSuccess)
            rc = B64_ERR_HIP;
        (void)hipStreamDestroy(cap);
        B64_TRY(rc);
        plan->use_graph = true;
    }
    *out = guard.release();
    return B64_OK;
}

extern "C" {

int b64_batch_plan_enqueue(b64_batch_plan *plan, void *stream) {
    if (!plan) return B64_ERR_INVALID_ARGS;
    hipStream_t s = as_stream(stream);
    // h_roots is coherent pinned host memory: the roots are written there by a kernel
    if (plan->single) return job_enqueue(&plan->jobs[0], s, plan->h_roots, true);
    if (plan->use_graph)
        HIP_TRY(hipGraphLaunch(plan->exec, s));
    else
        B64_TRY(plan->enqueue(s));
    return B64_OK;
}

int b64_batch_plan_collect(b64_batch_plan *plan) {
    if (!plan) return B64_ERR_INVALID_ARGS;
    size_t at = 0;
    for (size_t j = 0; j < plan->jobs.size(); j++) {
        memcpy(plan->jobs[j].roots_out, plan->h_roots + 32 * at, 32ull * plan->roots_per_job[j]);
        at += plan->roots_per_job[j];
    }
    return B64_OK;
}

int b64_batch_plan_run(b64_batch_plan *plan, void *stream) {
    B64_TRY(b64_batch_plan_enqueue(plan, stream));
    HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return b64_batch_plan_collect(plan);
}

uint32_t b64_batch_plan_n_launches(const b64_batch_plan *plan) { return plan ? (uint32_t)plan->launches.size() : 0; }

void b64_batch_plan_destroy(b64_batch_plan *plan) { delete plan; }

}  // extern "C"

// ===========================================================================
// Batched proofs: each job's FRI commit phase driven by its own transcript.
// Phase 0 commits every codeword (NTT passes, 16-element leaves, tree 0); each
// fold phase folds and commits the next oracle of every job that has one due,
// after the host has written the previous roots to the jobs' transcripts and
// sampled their challenges; a job whose rounds are over samples its query
// indices and writes its whole query-phase tape in one launch shared with the
// other jobs of that phase.
// Every phase is enqueued without waiting; the driver (b64_batch_prove, or the
// server's worker) waits for it and calls step() for the host part.
// ===========================================================================
namespace {

struct ProveJob {
    b64_prove_job jd;
    Layout L;
    uint32_t curr_round = 0, next_commit = 0;
    bool has_next = true;
    int k = 0;           // next oracle to commit
    int n_roots = 0;
    bool due = false;    // oracle k's fold challenges are in `unprocessed`
    std::vector<uint64_t> unprocessed;  // challenges as (lo, hi) pairs
    b64::FriQuery q;
    b64::QueryArgs qa;  // its query phase (out_at: granule offset in the phase's output)
    uint32_t n_rounds() const { return L.si->p.log_dim + L.si->p.log_batch_size; }

    // FRIFoldProver rounds (fold.rs:107-175): sample and receive challenges
    // until a commitment is due or every round has its challenge.
    void advance() {
        for (;;) {
            if (has_next && next_commit == curr_round) {
                due = true;
                return;
            }
            if (curr_round >= n_rounds()) return;
            uint8_t c[16];
            b64_transcript_sample(jd.transcript, c, 16);
            uint64_t w[2];
            memcpy(w, c, 16);
            unprocessed.push_back(w[0]);
            unprocessed.push_back(w[1]);
            curr_round++;
        }
    }
    // The root of the codeword or of the oracle just committed, written to the transcript.
    void take_root(const uint8_t *root) {
        b64_transcript_message(jd.transcript, root, 32);
        if (jd.roots_out) memcpy(jd.roots_out + 32 * n_roots, root, 32);
        n_roots++;
    }
};

// Device buffers grow in stream order (no device-wide synchronisation, so a
// serving loop's other streams keep running).
// Capacities at least double, so a buffer is reallocated O(log size) times.
inline size_t grown(size_t cap, size_t need) { return need > 2 * cap ? need : 2 * cap; }

int grow_device(uint8_t **d, size_t *cap, size_t need, hipStream_t s) {
    if (need <= *cap) return B64_OK;
    const size_t n = grown(*cap, need);
    if (*d) HIP_TRY(hipFreeAsync(*d, s));
    *d = nullptr;
    *cap = 0;
    HIP_TRY(hipMallocAsync((void **)d, n, s));
    *cap = n;
    return B64_OK;
}

int grow_pinned(uint8_t **h, size_t *cap, size_t need, unsigned flags = hipHostMallocDefault) {
    if (need <= *cap) return B64_OK;
    const size_t n = grown(*cap, need);
    if (*h) HIP_TRY(hipHostFree(*h));
    *h = nullptr;
    *cap = 0;
    HIP_TRY(hipHostMalloc((void **)h, n, flags));
    *cap = n;
    return B64_OK;
}

}  // namespace

// Buffers of a batch's phases; they only grow, so a caller that keeps one set
// (the server, per stream slot) stops allocating once it has seen its largest
// batch.  Only touched between phases, after the previous phase has completed.
// The pinned stage and output buffers are coherent host memory: kernels read
// the stage tables and write the roots there directly.
struct b64::ProveBuffers {
    uint8_t *d_tables = nullptr, *h_stage = nullptr;  // stage tables (and gather sources)
    uint8_t *d_out = nullptr;                         // query tapes
    uint8_t *h_out[2] = {};                           // alternate phases' roots and query tapes
    size_t d_tables_cap = 0, h_stage_cap = 0, d_out_cap = 0, h_out_cap[2] = {};
    ~ProveBuffers() {
        if (d_tables) (void)hipFree(d_tables);
        if (d_out) (void)hipFree(d_out);
        if (h_stage) (void)hipHostFree(h_stage);
        for (uint8_t *h : h_out)
            if (h) (void)hipHostFree(h);
    }
};

// Each phase carries, for every job, the next step it is ready for: the
// round that commits its next oracle (jobs grouped into multi-job launches),
// or, once its rounds are over, its query phase.  A job is complete one phase
// after its last round, whatever the other jobs still have to do.
struct b64::ProveBatch {
    enum class St { Rounds, Query, Done };
    std::vector<ProveJob> pj;
    std::vector<St> st;
    b64::ProveBuffers *buf = nullptr;
    bool first = true;                  // the next round is phase 0 (the codewords' commitments)
    std::vector<uint32_t> in_round, in_query;  // jobs in the enqueued phase
    std::vector<uint8_t> host;          // tables of the phase's launches
    std::vector<Launch> launches;
    size_t roots_off = 0;
    size_t q_args_off = 0, q_start_off = 0, q_idx_off = 0;  // the query launch's tables
    uint32_t q_blocks = 0;
    uint64_t n_gran = 0;                // granules of the phase's query tapes
    int ob = 0;                         // h_out buffer of the enqueued phase
    // query tapes of the completed phase, not yet decommitted: (job, byte offset in h_out[pending_ob])
    std::vector<std::pair<uint32_t, size_t>> pending;
    int pending_ob = 0;

    // The next phase from each job's state.
    int plan() {
        SlotMap slots;
        std::vector<const uint4 *> roots;
        std::vector<b64::QueryArgs> qargs;
        std::vector<uint32_t> qstart;
        std::vector<uint64_t> qidx;
        in_round.clear();
        in_query.clear();
        n_gran = 0;
        q_blocks = 0;
        for (uint32_t j = 0; j < pj.size(); j++) {
            ProveJob &J = pj[j];
            const Layout &L = J.L;
            uint8_t *ws = (uint8_t *)J.jd.d_workspace;
            if (st[j] != St::Rounds) continue;
            std::vector<Stage> stg;
            if (first) {
                B64_TRY(ntt_stages(L, J.jd.d_msg, J.jd.log_msg_len, (uint4 *)(ws + L.cw), stg));
                leaf16_stage(L, ws, nullptr, stg);
                B64_TRY(tree_stages((uint4 *)(ws + L.tree0), 1ull << L.rs, SLOT_TREE0, SLOT_ORACLE - 1, stg, roots));
            } else if (J.due) {
                const b64_fri_params &p = L.si->p;
                if (J.k >= L.n_oracles) return B64_ERR_INVALID_ARGS;
                const size_t n_ch = J.unprocessed.size() / 2;
                if (J.k == 0) {
                    // fold_interleaved (fold.rs:260-282) of the codeword into oracle 0
                    if (n_ch != (size_t)L.lbs || L.lbs != 4) return B64_ERR_INVALID_ARGS;
                    Stage f;
                    f.slot = SLOT_ORACLE - 1;
                    f.kind = b64::MultiKind::FoldInterleaved;
                    f.foldil.cw = (const uint4 *)(ws + L.cw);
                    f.foldil.out = (uint4 *)(ws + L.oracle[0]);
                    f.foldil.n_out = 1ull << L.rs;
                    for (int t = 0; t < 4; t++) f.foldil.chal[t] = to_u4(J.unprocessed.data() + 2 * t);
                    f.blocks = blocks_for(f.foldil.n_out);
                    stg.push_back(f);
                } else if (n_ch != p.fold_arities[J.k - 1]) {
                    return B64_ERR_INVALID_ARGS;
                }
                B64_TRY(oracle_stages(L, ws, J.k, J.unprocessed.data(), stg, roots));
            } else {
                // every commitment made: the query phase (finish_proof)
                if (J.curr_round != J.n_rounds() || J.k != L.n_oracles) return B64_ERR_EARLY_PROVER_FINISH;
                const b64::FriOracleView cw = {(const uint4 *)(ws + L.cw), (const uint4 *)(ws + L.tree0),
                                               L.rs + L.lbs, L.rs};
                std::vector<b64::FriOracleView> committed;
                for (int k = 0; k < L.n_oracles; k++)
                    committed.push_back({(const uint4 *)(ws + L.oracle[k]), (const uint4 *)(ws + L.tree[k]),
                                         L.olen[k], L.olen[k] - L.ocoset[k]});
                B64_TRY(J.q.prepare(L.si->p, cw, committed.data(), committed.size()));
                B64_TRY(J.q.sample(J.jd.transcript));
                B64_TRY(J.q.args(&J.qa));
                J.qa.idx_at = qidx.size();
                J.qa.out_at = n_gran;
                const uint32_t nb = b64::fri_query_blocks(J.qa);
                if ((uint64_t)q_blocks + nb >= (1ull << 31)) return B64_ERR_INVALID_ARGS;
                qstart.push_back(q_blocks);
                q_blocks += nb;
                n_gran += J.qa.n_gran;
                qidx.insert(qidx.end(), J.q.idx.begin(), J.q.idx.end());
                qargs.push_back(J.qa);
                st[j] = St::Query;
                in_query.push_back(j);
                continue;
            }
            for (const Stage &x : stg) slots[{x.slot, (int)x.kind}].push_back(x);
            in_round.push_back(j);
        }
        first = false;
        host.clear();
        launches.clear();
        B64_TRY(build_launches(slots, host, launches));
        roots_off = put_bytes(host, roots.data(), roots.size() * sizeof(void *));
        if (!in_query.empty()) {
            q_args_off = put_bytes(host, qargs.data(), qargs.size() * sizeof(b64::QueryArgs));
            q_start_off = put_bytes(host, qstart.data(), qstart.size() * sizeof(uint32_t));
            q_idx_off = put_bytes(host, qidx.data(), qidx.size() * sizeof(uint64_t));
        }
        return B64_OK;
    }

    // Output layout of a phase: the round's roots, then the query tapes.
    size_t granules_off() const { return align_up(32 * in_round.size()); }

    int enqueue(hipStream_t s) {
        const size_t n_roots = in_round.size(), out_bytes = granules_off() + 16 * n_gran;
        if (!host.empty()) {
            const size_t up = (host.size() + 15) & ~(size_t)15;
            B64_TRY(grow_device(&buf->d_tables, &buf->d_tables_cap, up, s));
            B64_TRY(grow_pinned(&buf->h_stage, &buf->h_stage_cap, up, hipHostMallocCoherent));
            memcpy(buf->h_stage, host.data(), host.size());
            B64_TRY(upload_tables(buf->h_stage, buf->d_tables, host.size(), s));
        }
        B64_TRY(grow_device(&buf->d_out, &buf->d_out_cap, 16 * n_gran, s));
        B64_TRY(grow_pinned(&buf->h_out[ob], &buf->h_out_cap[ob], out_bytes, hipHostMallocCoherent));
        if (n_roots) {
            B64_TRY(enqueue_launches(launches, buf->d_tables, s));
            // the roots go straight to the (coherent) host buffer
            k_gather_roots<<<(unsigned)((n_roots + 63) / 64), 64, 0, s>>>(
                (const uint4 *const *)(buf->d_tables + roots_off), (uint4 *)buf->h_out[ob], (int)n_roots);
            HIP_TRY(hipGetLastError());
        }
        if (!in_query.empty()) {
            HIP_TRY(b64::launch_fri_query((const b64::QueryArgs *)(buf->d_tables + q_args_off),
                                          (const uint32_t *)(buf->d_tables + q_start_off), (int)in_query.size(),
                                          q_blocks, (const uint64_t *)(buf->d_tables + q_idx_off),
                                          (uint4 *)buf->d_out, s));
            HIP_TRY(hipMemcpyAsync(buf->h_out[ob] + granules_off(), buf->d_out, 16 * n_gran, hipMemcpyDeviceToHost,
                                   s));
        }
        return B64_OK;
    }

    // The host part after the enqueued phase has completed: roots into the
    // transcripts and challenges sampled, then the next phase planned (*more
    // when there is one).  The phase's query tapes stay pending in its output
    // buffer, which the next phase does not write, until drain().
    int step(bool *more) {
        drain(nullptr);  // a caller that skipped it
        const uint8_t *h = buf->h_out[ob];
        for (size_t i = 0; i < in_round.size(); i++) {
            ProveJob &J = pj[in_round[i]];
            if (J.due) {
                const b64_fri_params &p = J.L.si->p;
                J.unprocessed.clear();
                J.due = false;
                if ((uint32_t)J.k < p.n_fold_arities)
                    J.next_commit += p.fold_arities[J.k];
                else
                    J.has_next = false;
                J.k++;
            }
            J.take_root(h + 32 * i);
            J.advance();
        }
        for (uint32_t j : in_query) pending.push_back({j, granules_off() + 16 * pj[j].qa.out_at});
        pending_ob = ob;
        B64_TRY(plan());
        ob ^= 1;
        *more = !in_round.empty() || !in_query.empty();
        return B64_OK;
    }

    // Decommits the pending query tapes: those jobs' proofs are complete
    // (appended to `finished`, indices into the batch, if not null).
    void drain(std::vector<uint32_t> *finished) {
        for (const auto &pq : pending) {
            const uint32_t j = pq.first;
            b64_transcript_decommit(pj[j].jd.transcript, buf->h_out[pending_ob] + pq.second, 16 * pj[j].qa.n_gran);
            st[j] = St::Done;
            if (finished) finished->push_back(j);
        }
        pending.clear();
    }
};

int b64::prove_batch_create(const b64_prove_job *jobs, uint32_t n_jobs, b64::ProveBuffers *buf,
                            b64::ProveBatch **out) {
    if (!jobs || !n_jobs || !buf || !out) return B64_ERR_INVALID_ARGS;
    std::unique_ptr<b64::ProveBatch> pb(new (std::nothrow) b64::ProveBatch());
    if (!pb) return B64_ERR_OUT_OF_MEMORY;
    pb->buf = buf;
    pb->pj.resize(n_jobs);
    for (uint32_t j = 0; j < n_jobs; j++) {
        const b64_prove_job &jd = jobs[j];
        if (!jd.d_msg || !jd.d_workspace || !jd.transcript) return B64_ERR_INVALID_ARGS;
        SizeInfo *si = nullptr;
        B64_TRY(size_info(jd.log_msg_len, &si));
        B64_TRY(size_device(si));
        pb->pj[j].jd = jd;
        layout_of(si, pb->pj[j].L);
        pb->pj[j].next_commit = si->p.log_batch_size;  // FRIFoldProver::new (fold.rs:70)
    }
    pb->st.assign(n_jobs, b64::ProveBatch::St::Rounds);
    B64_TRY(pb->plan());
    *out = pb.release();
    return B64_OK;
}

int b64::prove_batch_enqueue(b64::ProveBatch *pb, hipStream_t s) { return pb->enqueue(s); }
int b64::prove_batch_step(b64::ProveBatch *pb, bool *more) { return pb->step(more); }
void b64::prove_batch_drain(b64::ProveBatch *pb, std::vector<uint32_t> *finished) { pb->drain(finished); }
void b64::prove_batch_destroy(b64::ProveBatch *pb) { delete pb; }
b64::ProveBuffers *b64::prove_buffers_create(size_t reserve, hipStream_t s) {
    std::unique_ptr<b64::ProveBuffers> b(new (std::nothrow) b64::ProveBuffers());
    if (!b) return nullptr;
    if (reserve &&
        (grow_device(&b->d_tables, &b->d_tables_cap, reserve, s) != B64_OK ||
         grow_pinned(&b->h_stage, &b->h_stage_cap, reserve, hipHostMallocCoherent) != B64_OK ||
         grow_device(&b->d_out, &b->d_out_cap, 2 * reserve, s) != B64_OK ||
         grow_pinned(&b->h_out[0], &b->h_out_cap[0], 2 * reserve, hipHostMallocCoherent) != B64_OK ||
         grow_pinned(&b->h_out[1], &b->h_out_cap[1], 2 * reserve, hipHostMallocCoherent) != B64_OK))
        return nullptr;
    return b.release();
}
void b64::prove_buffers_destroy(b64::ProveBuffers *b) { delete b; }

namespace {

// Phase buffers of b64_batch_prove, kept per device for later calls (pinned host
// allocations cost far more than a small batch's work); a call takes a set
// from the pool and returns it when it ends, so concurrent calls get their own.
constexpr int POOL_DEVICES = 64;
std::mutex g_prove_pool_mu;
std::vector<b64::ProveBuffers *> g_prove_pool[POOL_DEVICES];

struct PooledBuffers {
    int dev = -1;
    b64::ProveBuffers *buf = nullptr;
    int acquire() {
        HIP_TRY(hipGetDevice(&dev));
        if (dev < 0 || dev >= POOL_DEVICES) return B64_ERR_NO_DEVICE;
        {
            std::lock_guard<std::mutex> g(g_prove_pool_mu);
            if (!g_prove_pool[dev].empty()) {
                buf = g_prove_pool[dev].back();
                g_prove_pool[dev].pop_back();
                return B64_OK;
            }
        }
        buf = b64::prove_buffers_create(0, nullptr);
        return buf ? B64_OK : B64_ERR_OUT_OF_MEMORY;
    }
    ~PooledBuffers() {
        if (!buf) return;
        std::lock_guard<std::mutex> g(g_prove_pool_mu);
        g_prove_pool[dev].push_back(buf);
    }
};

// Parts of a b64_batch_prove call, each with its own stream.
constexpr int PROVE_PARTS = 2;

// A side stream (non-blocking, default priority) and an event for a part of
// b64_batch_prove's jobs, pooled per device like the buffers.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
};
std::vector<SideStream> g_side_pool[POOL_DEVICES];

struct PooledSide {
    int dev = -1;
    SideStream side;
    bool held = false;
    int acquire() {
        HIP_TRY(hipGetDevice(&dev));
        if (dev < 0 || dev >= POOL_DEVICES) return B64_ERR_NO_DEVICE;
        {
            std::lock_guard<std::mutex> g(g_prove_pool_mu);
            if (!g_side_pool[dev].empty()) {
                side = g_side_pool[dev].back();
                g_side_pool[dev].pop_back();
                held = true;
                return B64_OK;
            }
        }
        HIP_TRY(hipStreamCreateWithFlags(&side.s, hipStreamNonBlocking));
        if (hipEventCreateWithFlags(&side.ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(side.s);
            return B64_ERR_HIP;
        }
        held = true;
        return B64_OK;
    }
    ~PooledSide() {
        if (!held) return;
        std::lock_guard<std::mutex> g(g_prove_pool_mu);
        g_side_pool[dev].push_back(side);
    }
};

// One batch's phases on one stream: the host part of a completed phase, the
// next phase enqueued, then the completed phase's query tapes decommitted.
struct PhaseDriver {
    b64::ProveBatch *pb;
    hipStream_t s;
    bool more = true;
    int turn() {
        if (hipStreamSynchronize(s) != hipSuccess) return B64_ERR_HIP;
        B64_TRY(pb->step(&more));
        const int rc = more ? pb->enqueue(s) : B64_OK;
        pb->drain(nullptr);
        return rc;
    }
};

}  // namespace

extern "C" int b64_batch_prove(const b64_prove_job *jobs, uint32_t n_jobs, void *stream) {
    if (!jobs || !n_jobs) return B64_ERR_INVALID_ARGS;
    for (uint32_t j = 0; j < n_jobs; j++)
        if (!jobs[j].d_msg || !jobs[j].d_workspace || !jobs[j].transcript) return B64_ERR_INVALID_ARGS;
    const hipStream_t s = as_stream(stream);
    // The jobs in PROVE_PARTS parts (dealt round-robin in size order, so every
    // part has work in every phase), each on a stream of its own: the host part
    // of one part's phase runs while the device works on the others'.
    const int parts = n_jobs < (uint32_t)PROVE_PARTS ? (int)n_jobs : PROVE_PARTS;
    std::vector<uint32_t> order(n_jobs);
    for (uint32_t j = 0; j < n_jobs; j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return jobs[a].log_msg_len > jobs[b].log_msg_len; });
    std::vector<b64_prove_job> part[PROVE_PARTS];
    for (uint32_t i = 0; i < n_jobs; i++) part[i % parts].push_back(jobs[order[i]]);
    PooledBuffers pool[PROVE_PARTS];
    PooledSide side[PROVE_PARTS];
    hipStream_t st[PROVE_PARTS];
    for (int h = 0; h < parts; h++) {
        B64_TRY(pool[h].acquire());
        st[h] = s;
        if (h > 0) {
            // a side stream starts after the work the caller queued on s (the messages)
            B64_TRY(side[h].acquire());
            st[h] = side[h].side.s;
            HIP_TRY(hipEventRecord(side[h].side.ev, s));
            HIP_TRY(hipStreamWaitEvent(st[h], side[h].side.ev, 0));
        }
    }
    std::unique_ptr<b64::ProveBatch> pb[PROVE_PARTS];
    PhaseDriver drv[PROVE_PARTS];
    for (int h = 0; h < parts; h++) {
        b64::ProveBatch *raw = nullptr;
        B64_TRY(b64::prove_batch_create(part[h].data(), (uint32_t)part[h].size(), pool[h].buf, &raw));
        pb[h].reset(raw);
        drv[h] = {raw, st[h]};
    }
    int rc = B64_OK;
    for (int h = 0; h < parts && rc == B64_OK; h++) rc = pb[h]->enqueue(st[h]);
    for (bool more = true; rc == B64_OK && more;) {
        more = false;
        for (int h = 0; h < parts && rc == B64_OK; h++)
            if (drv[h].more) {
                rc = drv[h].turn();
                more = more || drv[h].more;
            }
    }
    // the buffers go back to the pool only once nothing of this call is in flight
    if (rc != B64_OK)
        for (int h = 0; h < parts; h++) (void)hipStreamSynchronize(st[h]);
    return rc;
}
