// Circuit planner: exact per-gate sweeps, or fused LDS tile passes.
//
// The fused lowering works in three steps:
//   1. peephole merging on the gate list: runs of single-qubit gates on one
//      qubit become one 2x2 matrix; the CNOT(c,t) . diag(t) . CNOT(c,t)
//      sandwich that the reference's QAOA builder emits for every cost term
//      (src/algorithms.rs:122-126) becomes one two-qubit diagonal; CZ becomes
//      a diagonal;
//   2. greedy tile selection: starting from the contiguous low qubits, add the
//      qubit that lets the most pending gates run, until the tile holds M
//      qubits (gates run in per-qubit dependency order, so any order the
//      schedule picks is equivalent to the circuit order);
//   3. emission: each tile becomes one pass, i.e. one HBM sweep.
#include "qf_plan.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>

namespace qf {

namespace {

inline cd mul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cd add(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }

// Product r = x * y of 2x2 matrices stored (m00, m01, m10, m11).
void mat2_mul(const cd x[4], const cd y[4], cd r[4]) {
    r[0] = add(mul(x[0], y[0]), mul(x[1], y[2]));
    r[1] = add(mul(x[0], y[1]), mul(x[1], y[3]));
    r[2] = add(mul(x[2], y[0]), mul(x[3], y[2]));
    r[3] = add(mul(x[2], y[1]), mul(x[3], y[3]));
}

bool is_zero(cd a) { return a.re == 0.0 && a.im == 0.0; }

enum NodeKind { N_MAT1, N_DIAG1, N_DIAG2, N_CNOT, N_SWAP, N_ROT, N_ROTZZ, N_MAT2 };

struct Node {
    NodeKind kind;
    uint32_t q[2];
    int nq;
    cd m[4];  // MAT1 matrix, DIAG1 (d0, d1), DIAG2 d[bit(q1)*2 + bit(q0)]
    cd m4[16];  // MAT2: row-major 4x4 over bit(q1)*2 + bit(q0)
    int axis = 0, param = 0, slot = 0, flags = 0;  // N_ROT / N_ROTZZ
    bool dead = false;
    int prev[2] = {-1, -1};  // previous live node on q[0], q[1]
    bool diag() const {
        return kind == N_DIAG1 || kind == N_DIAG2 || kind == N_ROTZZ || (kind == N_ROT && axis == 2);
    }
};

int slot_of(const Node& n, uint32_t q) { return n.q[0] == q ? 0 : 1; }

// Step 1: peephole merging.
std::vector<Node> lower(const std::vector<GateIn>& gates, uint32_t qubits) {
    std::vector<Node> nodes;
    nodes.reserve(gates.size());
    std::vector<int> last(qubits, -1);
    auto push = [&](Node n) {
        for (int k = 0; k < n.nq; ++k) n.prev[k] = last[n.q[k]];
        nodes.push_back(n);
        const int id = int(nodes.size()) - 1;
        for (int k = 0; k < n.nq; ++k) last[n.q[k]] = id;
    };
    for (const GateIn& g : gates) {
        if (g.kind == 0) {
            Node n{};
            n.nq = 1;
            n.q[0] = g.a;
            std::memcpy(n.m, g.m, sizeof(n.m));
            n.kind = (is_zero(g.m[1]) && is_zero(g.m[2])) ? N_DIAG1 : N_MAT1;
            const int L = last[g.a];
            if (L >= 0 && (nodes[L].kind == N_MAT1 || nodes[L].kind == N_DIAG1)) {
                cd r[4];
                mat2_mul(n.m, nodes[L].m, r);  // new gate after the old one
                std::memcpy(nodes[L].m, r, sizeof(r));
                nodes[L].kind = (is_zero(r[1]) && is_zero(r[2])) ? N_DIAG1 : N_MAT1;
                continue;
            }
            if (L >= 0 && n.kind == N_DIAG1 && nodes[L].kind == N_DIAG2) {
                Node& D = nodes[L];
                const int s = slot_of(D, g.a);
                for (int k = 0; k < 4; ++k) D.m[k] = mul(((k >> s) & 1) ? n.m[3] : n.m[0], D.m[k]);
                continue;
            }
            push(n);
        } else if (g.kind == 6) {  // dense two-qubit matrix
            Node n{};
            n.kind = N_MAT2;
            n.nq = 2;
            n.q[0] = g.a;
            n.q[1] = g.b;
            std::memcpy(n.m4, g.m4, sizeof(n.m4));
            push(n);
        } else if (g.kind == 5) {  // rotation with a per-instance angle
            Node n{};
            n.kind = N_ROT;
            n.nq = 1;
            n.q[0] = g.a;
            n.axis = g.axis;
            n.param = g.param;
            n.slot = g.slot;
            n.flags = g.flags;
            push(n);
        } else if (g.kind == 1) {  // CNOT(a control, b target)
            const uint32_t c = g.a, t = g.b;
            const int Lt = last[t];
            const bool mid_diag = Lt >= 0 && (nodes[Lt].kind == N_DIAG1 || (nodes[Lt].kind == N_ROT && nodes[Lt].axis == 2));
            if (mid_diag) {
                const int K = nodes[Lt].prev[0];
                if (K >= 0 && nodes[K].kind == N_CNOT && nodes[K].q[0] == c && nodes[K].q[1] == t && last[c] == K) {
                    Node& D = nodes[K];
                    if (nodes[Lt].kind == N_DIAG1) {
                        // CNOT(c,t) diag(d0,d1)_t CNOT(c,t) = diag over (c,t): c=0 -> d(t), c=1 -> d(!t).
                        const cd d0 = nodes[Lt].m[0], d1 = nodes[Lt].m[3];
                        D.kind = N_DIAG2;
                        // slots: q[0]=c, q[1]=t ; index = bit(t)*2 + bit(c)
                        D.m[0] = d0;  // c0 t0
                        D.m[1] = d1;  // c1 t0 -> X d X on t: d1
                        D.m[2] = d1;  // c0 t1
                        D.m[3] = d0;  // c1 t1
                    } else {
                        // CNOT(c,t) Rz(theta)_t CNOT(c,t) = exp(-i theta Z_c Z_t / 2); its
                        // generator at this point is Z_c Z_t (the QAOA cost term).
                        D.kind = N_ROTZZ;
                        D.axis = 2;
                        D.param = nodes[Lt].param;
                        D.slot = nodes[Lt].slot;
                        D.flags = nodes[Lt].flags;
                    }
                    nodes[Lt].dead = true;
                    last[t] = K;
                    continue;
                }
            }
            Node n{};
            n.kind = N_CNOT;
            n.nq = 2;
            n.q[0] = c;
            n.q[1] = t;
            push(n);
        } else if (g.kind == 2 || g.kind == 4) {  // CZ or general two-qubit diagonal
            Node n{};
            n.kind = N_DIAG2;
            n.nq = 2;
            n.q[0] = g.a;
            n.q[1] = g.b;
            if (g.kind == 2) {
                n.m[0] = n.m[1] = n.m[2] = cd{1, 0};
                n.m[3] = cd{-1, 0};
            } else {
                std::memcpy(n.m, g.m, sizeof(n.m));
            }
            const int La = last[g.a], Lb = last[g.b];
            if (La >= 0 && La == Lb && nodes[La].kind == N_DIAG2) {  // diag2 . diag2 on the same pair
                Node& D = nodes[La];
                const bool same = D.q[0] == g.a;
                for (int k = 0; k < 4; ++k) {
                    const int kk = same ? k : (((k & 1) << 1) | (k >> 1));
                    D.m[kk] = mul(n.m[k], D.m[kk]);
                }
                continue;
            }
            push(n);
        } else {  // SWAP
            Node n{};
            n.kind = N_SWAP;
            n.nq = 2;
            n.q[0] = g.a;
            n.q[1] = g.b;
            push(n);
        }
    }
    std::vector<Node> live;
    for (auto& n : nodes)
        if (!n.dead) live.push_back(n);
    return live;
}

// Arithmetic of a node per amplitude, in half-instruction units, for
// balancing passes: run-time X/Y rotations and literal 1-qubit matrices are
// a multiply-add per output component, Z rotations and other diagonals a
// per-register factor, CNOT/SWAP renamings or selects.
int node_cost(const Node& n) {
    switch (n.kind) {
        case N_ROT: return n.axis == 2 ? 4 : 8;
        case N_MAT1: return 6;
        case N_MAT2: return 32;
        case N_CNOT:
        case N_SWAP: return 1;
        default: return 2;
    }
}

struct Scheduler {
    const std::vector<Node>& nodes;
    uint32_t qubits;
    std::vector<std::vector<int>> queue;  // node ids per qubit, circuit order
    std::vector<size_t> head;
    size_t remaining;
    // 0: every qubit of a non-diagonal gate must be in the set;
    // 1: a CNOT needs only its target (the control bit is read from the index).
    int mode = 0;
    // Tie break among qubits that let equally many gates run (choose_set):
    // 0: most pending gates; 1: earliest pending gate; 2: earliest pending
    // gate counted in blocks of `block` gates, then most pending gates.
    int tie = 0;
    int block = 1;
    // When positive, a pass that is committed (simulate with an order) stops
    // taking nodes once their node_cost sum would exceed cap.
    int cap = 0;

    Scheduler(const std::vector<Node>& n, uint32_t nq, int m = 0)
        : nodes(n), qubits(nq), queue(nq), head(nq, 0), mode(m) {
        for (size_t i = 0; i < nodes.size(); ++i)
            for (int k = 0; k < nodes[i].nq; ++k) queue[nodes[i].q[k]].push_back(int(i));
        remaining = nodes.size();
    }

    bool at_head(const std::vector<size_t>& h, int id) const {
        const Node& n = nodes[id];
        for (int k = 0; k < n.nq; ++k) {
            const auto& qq = queue[n.q[k]];
            if (h[n.q[k]] >= qq.size() || qq[h[n.q[k]]] != id) return false;
        }
        return true;
    }
    bool fits(const Node& n, const std::vector<char>& in) const {
        if (n.diag()) return true;
        if (mode == 1 && n.kind == N_CNOT) return in[n.q[1]] != 0;
        for (int k = 0; k < n.nq; ++k)
            if (!in[n.q[k]]) return false;
        return true;
    }

    // Runs every node that can run with tile set `in`, from heads `h`
    // (updated).  Records the order in `out` when given.  Returns the weighted
    // count (non-diagonal nodes weigh 2, diagonal 1).
    int simulate(std::vector<size_t>& h, const std::vector<char>& in, std::vector<int>* out) const {
        int score = 0, cost = 0;
        std::deque<int> work;
        for (uint32_t q = 0; q < qubits; ++q)
            if (h[q] < queue[q].size()) work.push_back(queue[q][h[q]]);
        while (!work.empty()) {
            const int id = work.front();
            work.pop_front();
            if (!at_head(h, id)) continue;
            const Node& n = nodes[id];
            if (!fits(n, in)) continue;
            if (out && cap > 0 && cost > 0 && cost + node_cost(n) > cap) break;
            cost += node_cost(n);
            for (int k = 0; k < n.nq; ++k) ++h[n.q[k]];
            score += n.diag() ? 1 : 2;
            if (out) out->push_back(id);
            for (int k = 0; k < n.nq; ++k) {
                const uint32_t q = n.q[k];
                if (h[q] < queue[q].size()) work.push_back(queue[q][h[q]]);
            }
        }
        return score;
    }
};

// Greedy set selection: starting from `in` (size `size`), add candidates
// until the set holds `target` qubits, each time taking the qubit that lets
// the most pending gates run; when no single qubit helps, seed the set with
// the qubits of the earliest runnable gate.  Returns false if nothing can run.
bool choose_set(const Scheduler& sch, const std::vector<Node>& nodes, std::vector<char>& in, uint32_t size,
                uint32_t target, const std::vector<uint32_t>& cand) {
    std::vector<size_t> h_base = sch.head;
    int base_score = sch.simulate(h_base, in, nullptr);
    while (size < target) {
        int best = -1, best_score = -1, best_tie = std::numeric_limits<int>::min();
        for (uint32_t q : cand) {
            if (in[q]) continue;
            in[q] = 1;
            std::vector<size_t> h = sch.head;
            const int sc = sch.simulate(h, in, nullptr);
            in[q] = 0;
            // Tie break (Scheduler::tie), then the lowest index.
            const bool pend = sch.head[q] < sch.queue[q].size();
            const int work = int(sch.queue[q].size() - sch.head[q]);
            const int first = pend ? sch.queue[q][sch.head[q]] : int(nodes.size());
            int tie = work;
            if (sch.tie == 1) tie = -first;
            if (sch.tie == 2) tie = (int(nodes.size()) - first / sch.block) * 64 + std::min(63, work);
            if (sc > best_score || (sc == best_score && tie > best_tie)) {
                best = int(q);
                best_score = sc;
                best_tie = tie;
            }
        }
        if (best < 0) break;
        if (best_score <= base_score) {
            int pick = -1;
            for (uint32_t q : cand) {
                if (h_base[q] >= sch.queue[q].size()) continue;
                const int id = sch.queue[q][h_base[q]];
                if (!sch.at_head(h_base, id)) continue;
                const Node& nd = nodes[id];
                if (nd.diag()) continue;
                uint32_t missing = 0;
                bool ok = true;
                for (int k = 0; k < nd.nq; ++k) {
                    const bool needed = !(sch.mode == 1 && nd.kind == N_CNOT && k == 0);
                    if (!needed || in[nd.q[k]]) continue;
                    if (std::find(cand.begin(), cand.end(), nd.q[k]) == cand.end()) ok = false;
                    ++missing;
                }
                if (!ok || missing == 0 || missing > target - size) continue;
                if (pick < 0 || id < pick) pick = id;
            }
            if (pick >= 0) {
                const Node& nd = nodes[pick];
                for (int k = 0; k < nd.nq; ++k) {
                    const bool needed = !(sch.mode == 1 && nd.kind == N_CNOT && k == 0);
                    if (needed && !in[nd.q[k]]) {
                        in[nd.q[k]] = 1;
                        ++size;
                    }
                }
                h_base = sch.head;
                base_score = sch.simulate(h_base, in, nullptr);
                continue;
            }
        }
        in[best] = 1;
        ++size;
        h_base = sch.head;
        base_score = sch.simulate(h_base, in, nullptr);
    }
    std::vector<size_t> h = sch.head;
    return sch.simulate(h, in, nullptr) > 0;
}

// The set of a compiled pass with one pass of lookahead: candidates from the
// four tie-break policies, each scored by the gates it runs plus the gates
// the greedy set after it (same low-slot holders) would run.  The first best
// candidate wins.
bool choose_set_lookahead(Scheduler& sch, const std::vector<Node>& nodes, std::vector<char>& in, uint32_t low,
                          uint32_t M, const std::vector<uint32_t>& all_q, const std::vector<uint32_t>& at) {
    const int tie0 = sch.tie, block0 = sch.block;
    const int n = int(sch.qubits);
    const int pols[4][2] = {{0, 1}, {1, 1}, {2, n}, {2, n + n / 3}};
    const std::vector<size_t> head0 = sch.head;
    std::vector<char> best_in;
    long best = -1;
    for (const auto& pol : pols) {
        sch.tie = pol[0];
        sch.block = pol[1] > 0 ? pol[1] : 1;
        std::vector<char> cand(in);
        if (!choose_set(sch, nodes, cand, low, M, all_q)) continue;
        std::vector<size_t> h = head0;
        const int now = sch.simulate(h, cand, nullptr);
        sch.tie = tie0;
        sch.block = block0;
        sch.head = h;
        std::vector<char> nx(sch.qubits, 0);
        for (uint32_t b = 0; b < low; ++b) nx[at[b]] = 1;
        int next = 0;
        if (choose_set(sch, nodes, nx, low, M, all_q)) {
            std::vector<size_t> h2 = h;
            next = sch.simulate(h2, nx, nullptr);
        }
        sch.head = head0;
        if (long(now) + next > best) {
            best = long(now) + next;
            best_in = cand;
        }
    }
    sch.tie = tie0;
    sch.block = block0;
    if (best < 0) return false;
    in = best_in;
    return true;
}

TileOp make_reg_op(const Node& n, const std::vector<int>& reg_index);

// Splits one pass's ops into register phases of RB tile bits each.  Emits a
// TOP_PHASE header (packed tile bits of the register set, register index i
// <-> 4 bits at 4*i) before each phase's ops.
std::string build_phases_tie(const std::vector<Node>& nodes, const std::vector<int>& order, uint32_t qubits,
                             const std::vector<int>& tile_bit, uint32_t RB, std::vector<TileOp>& ops_out,
                             uint64_t& phase_count, int tie);

// Register phases of one pass: the greedy split under each tie-break policy
// of the set choice, keeping the one with the fewest phases (each phase
// change is an exchange of the whole tile through LDS).
std::string build_phases(const std::vector<Node>& nodes, const std::vector<int>& order, uint32_t qubits,
                         const std::vector<int>& tile_bit, uint32_t RB, std::vector<TileOp>& ops_out,
                         uint64_t& phase_count) {
    std::vector<TileOp> best;
    uint64_t best_n = 0;
    std::string err;
    for (int tie = 0; tie < 2; ++tie) {
        std::vector<TileOp> ops;
        uint64_t n = 0;
        std::string e = build_phases_tie(nodes, order, qubits, tile_bit, RB, ops, n, tie);
        if (!e.empty()) {
            err = e;
            continue;
        }
        if (best.empty() || n < best_n) {
            best = std::move(ops);
            best_n = n;
        }
    }
    if (best.empty()) return err;
    ops_out.insert(ops_out.end(), best.begin(), best.end());
    phase_count += best_n;
    return "";
}

std::string build_phases_tie(const std::vector<Node>& nodes, const std::vector<int>& order, uint32_t qubits,
                             const std::vector<int>& tile_bit, uint32_t RB, std::vector<TileOp>& ops_out,
                             uint64_t& phase_count, int tie) {
    std::vector<Node> sub;
    sub.reserve(order.size());
    for (int id : order) sub.push_back(nodes[id]);
    Scheduler ps(sub, qubits, 1);
    ps.tie = tie;
    std::vector<uint32_t> tq;
    for (uint32_t q = 0; q < qubits; ++q)
        if (tile_bit[q] >= 0) tq.push_back(q);
    size_t done = 0;
    while (done < sub.size()) {
        std::vector<char> in(qubits, 0);
        if (!choose_set(ps, sub, in, 0, RB, tq)) {
            // Only diagonal gates can be pending here with an empty useful set;
            // fill the set with any tile qubits so they still run.
            uint32_t size = 0;
            for (uint32_t q : tq)
                if (size < RB && !in[q]) {
                    in[q] = 1;
                    ++size;
                }
        }
        std::vector<int> ord;
        ps.simulate(ps.head, in, &ord);
        if (ord.empty()) return "phase planner made no progress";
        done += ord.size();
        std::vector<int> reg_index(qubits, -1);
        TileOp hdr;
        std::memset(&hdr, 0, sizeof(hdr));
        hdr.kind = TOP_PHASE;
        uint32_t i = 0, packed = 0;
        for (uint32_t q : tq)
            if (in[q] && i < RB) {
                reg_index[q] = int(i);
                packed |= uint32_t(tile_bit[q]) << (4 * i);
                ++i;
            }
        // Fill unused register slots with further tile bits (the set must have RB distinct bits).
        for (uint32_t q : tq)
            if (i < RB && reg_index[q] < 0) {
                reg_index[q] = int(i);
                packed |= uint32_t(tile_bit[q]) << (4 * i);
                ++i;
            }
        hdr.q0 = int32_t(packed);
        hdr.q1 = int32_t(RB);
        ops_out.push_back(hdr);
        ++phase_count;
        for (int id : ord) ops_out.push_back(make_reg_op(sub[id], reg_index));
    }
    return "";
}

TileOp make_tile_op(const Node& n, const std::vector<int>& tile_bit) {
    TileOp op;
    std::memset(&op, 0, sizeof(op));
    switch (n.kind) {
        case N_MAT1:
            op.kind = TOP_MAT1;
            op.q0 = tile_bit[n.q[0]];
            op.q1 = -1;
            std::memcpy(op.m, n.m, 4 * sizeof(cd));
            break;
        case N_DIAG1:
            op.kind = TOP_DIAG;
            op.q0 = int(n.q[0]);
            op.q1 = -1;
            op.m[0] = n.m[0].re; op.m[1] = n.m[0].im;
            op.m[2] = n.m[3].re; op.m[3] = n.m[3].im;
            op.m[4] = op.m[6] = 0.0; op.m[5] = op.m[7] = 0.0;
            break;
        case N_DIAG2:
            op.kind = TOP_DIAG;
            op.q0 = int(n.q[0]);
            op.q1 = int(n.q[1]);
            std::memcpy(op.m, n.m, 4 * sizeof(cd));
            break;
        case N_CNOT:
            op.kind = TOP_CNOT;
            op.q0 = tile_bit[n.q[0]];
            op.q1 = tile_bit[n.q[1]];
            break;
        case N_SWAP:
            op.kind = TOP_SWAP;
            op.q0 = tile_bit[n.q[0]];
            op.q1 = tile_bit[n.q[1]];
            break;
        case N_ROT:
            op.kind = TOP_ROT;
            op.q0 = n.axis == 2 ? int(n.q[0]) : tile_bit[n.q[0]];
            op.q1 = n.axis;
            op.flags = n.flags;
            op.param = n.param;
            op.slot = n.slot;
            break;
        case N_ROTZZ:
            op.kind = TOP_ROTZZ;
            op.q0 = int(n.q[0]);
            op.q1 = int(n.q[1]);
            op.flags = n.flags;
            op.param = n.param;
            op.slot = n.slot;
            break;
        case N_MAT2:
            op.kind = TOP_MAT2;
            op.q0 = tile_bit[n.q[0]];
            op.q1 = tile_bit[n.q[1]];
            std::memcpy(op.m, n.m4, sizeof(n.m4));
            break;
    }
    return op;
}

// A dense 4x4 over (bit(hi) * 2 + bit(lo)) re-indexed for the bits in the
// other order: swaps the two index bits of rows and columns.
void swap_mat2_bits(const cd in[16], cd out[16]) {
    auto sw = [](int k) { return ((k & 1) << 1) | (k >> 1); };
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[sw(r) * 4 + sw(c)] = in[r * 4 + c];
}

// Arithmetic form of a 2x2 matrix (see MAT_*).
int32_t mat_form(const cd m[4]) {
    if (m[0].re == 0 && m[0].im == 0 && m[3].re == 0 && m[3].im == 0 && m[1].re == 1 && m[1].im == 0 &&
        m[2].re == 1 && m[2].im == 0)
        return MAT_SWAP;
    if (m[0].im == 0 && m[1].im == 0 && m[2].im == 0 && m[3].im == 0) return MAT_REAL;
    if (m[0].im == 0 && m[3].im == 0 && m[1].re == 0 && m[2].re == 0) return MAT_CROSS;
    return MAT_GENERAL;
}

TileOp make_reg_op(const Node& n, const std::vector<int>& reg) {
    TileOp op;
    std::memset(&op, 0, sizeof(op));
    switch (n.kind) {
        case N_MAT1:
            op.kind = TOP_MAT1;
            op.q0 = reg[n.q[0]];
            op.q1 = -1;
            op.flags = mat_form(n.m);
            std::memcpy(op.m, n.m, 4 * sizeof(cd));
            break;
        case N_DIAG1:
            op.kind = TOP_DIAG;
            op.q0 = int(n.q[0]);
            op.q1 = -1;
            op.m[0] = n.m[0].re; op.m[1] = n.m[0].im;
            op.m[2] = n.m[3].re; op.m[3] = n.m[3].im;
            break;
        case N_DIAG2:
            op.kind = TOP_DIAG;
            op.q0 = int(n.q[0]);
            op.q1 = int(n.q[1]);
            std::memcpy(op.m, n.m, 4 * sizeof(cd));
            break;
        case N_CNOT:
            op.kind = TOP_CNOT;
            op.q0 = int(n.q[0]);      // control: global qubit
            op.q1 = reg[n.q[1]];      // target: register index
            break;
        case N_SWAP:
            op.kind = TOP_SWAP;
            op.q0 = reg[n.q[0]];
            op.q1 = reg[n.q[1]];
            break;
        case N_ROT:
            op.kind = TOP_ROT;
            op.q0 = n.axis == 2 ? int(n.q[0]) : reg[n.q[0]];
            op.q1 = n.axis;
            op.flags = n.flags;
            op.param = n.param;
            op.slot = n.slot;
            break;
        case N_ROTZZ:
            op.kind = TOP_ROTZZ;
            op.q0 = int(n.q[0]);
            op.q1 = int(n.q[1]);
            op.flags = n.flags;
            op.param = n.param;
            op.slot = n.slot;
            break;
        case N_MAT2: {
            // Register form: q0 < q1 (the kernels instantiate ordered pairs only).
            op.kind = TOP_MAT2;
            const int r0 = reg[n.q[0]], r1 = reg[n.q[1]];
            if (r0 < r1) {
                op.q0 = r0;
                op.q1 = r1;
                std::memcpy(op.m, n.m4, sizeof(n.m4));
            } else {
                op.q0 = r1;
                op.q1 = r0;
                cd t[16];
                swap_mat2_bits(n.m4, t);
                std::memcpy(op.m, t, sizeof(t));
            }
            break;
        }
    }
    return op;
}

}  // namespace

std::string describe_plan(const Plan& plan) {
    std::string s;
    char buf[64];
    auto num = [&](double v) {
        std::snprintf(buf, sizeof(buf), "%.17g", v);
        s += buf;
    };
    auto cplx = [&](const double* m, int count) {
        s += "[";
        for (int k = 0; k < 2 * count; ++k) {
            if (k) s += ",";
            num(m[k]);
        }
        s += "]";
    };
    s += "{\"qubits\":" + std::to_string(plan.qubits) + ",\"compiled\":" + std::to_string(plan.compiled) +
         ",\"relabeled\":" + std::to_string(plan.relabeled) +
         ",\"scaled_rotations\":" + std::to_string(plan.scaled_rotations) + ",\"passes\":[";
    for (size_t i = 0; i < plan.passes.size(); ++i) {
        const Pass& p = plan.passes[i];
        if (i) s += ",";
        if (p.kind == Pass::TILE) {
            s += "{\"kind\":\"tile\",\"tile_bits\":" + std::to_string(p.tile_bits) + ",\"reg_bits\":" +
                 std::to_string(p.reg_bits) + ",\"low\":" + std::to_string(p.low) + ",\"high\":[";
            for (size_t k = 0; k < p.high.size(); ++k) s += (k ? "," : "") + std::to_string(p.high[k]);
            s += "],\"store\":[";
            for (size_t k = 0; k < p.store_phys.size(); ++k) s += (k ? "," : "") + std::to_string(p.store_phys[k]);
            s += "],\"ops\":[";
            for (size_t k = 0; k < p.ops.size(); ++k) {
                const TileOp& o = p.ops[k];
                if (k) s += ",";
                s += "{\"kind\":" + std::to_string(o.kind) + ",\"q0\":" + std::to_string(o.q0) + ",\"q1\":" +
                     std::to_string(o.q1) + ",\"flags\":" + std::to_string(o.flags) + ",\"slot\":" +
                     std::to_string(o.slot) + ",\"m\":";
                cplx(o.m, o.kind == TOP_MAT2 ? 16 : 4);
                s += "}";
            }
            s += "]}";
        } else if (p.kind == Pass::DENSE) {
            s += "{\"kind\":\"dense\",\"block\":[";
            for (size_t k = 0; k < p.block.size(); ++k) s += (k ? "," : "") + std::to_string(p.block[k]);
            s += "],\"U\":";
            cplx(&p.U[0].re, int(p.U.size()));
            for (int list = 0; list < 2; ++list) {
                const auto& L = list ? p.pre_blk : p.pre_col;
                s += list ? ",\"pre_blk\":[" : ",\"pre_col\":[";
                for (size_t k = 0; k < L.size(); ++k) {
                    if (k) s += ",";
                    s += "{\"a\":" + std::to_string(L[k].a) + ",\"b\":" + std::to_string(L[k].b) + ",\"d\":";
                    cplx(&L[k].d[0].re, 4);
                    s += "}";
                }
                s += "]";
            }
            s += "}";
        } else {
            static const char* names[] = {"sweep_1q", "sweep_2q", "tile", "sweep_diag"};
            s += std::string("{\"kind\":\"") + names[p.kind] + "\",\"op2\":" + std::to_string(p.op2) +
                 ",\"q0\":" + std::to_string(p.q0) + ",\"q1\":" + std::to_string(p.q1) + ",\"m\":";
            cplx(&p.m[0].re, 4);
            s += "}";
        }
    }
    s += "],\"stats\":{\"passes\":" + std::to_string(plan.stats.passes) +
         ",\"ops_in\":" + std::to_string(plan.stats.ops_in) +
         ",\"ops_applied\":" + std::to_string(plan.stats.ops_applied) +
         ",\"max_ops_per_pass\":" + std::to_string(plan.stats.max_ops_per_pass) +
         ",\"dense_blocks\":" + std::to_string(plan.stats.dense_blocks) +
         ",\"phases\":" + std::to_string(plan.stats.phases) + "}}";
    return s;
}

namespace {

// Relabeling of compiled passes: set QF_PLAN_RELABEL=0 to keep the identity layout.
bool relabel_enabled() {
    const char* e = std::getenv("QF_PLAN_RELABEL");
    return !(e && e[0] == '0');
}

// Low bits and their relabeling.  The low index bits 0..low-1 must be in every
// tile (they make the 128-byte runs of a tile), so whichever qubits hold them
// take tile slots in every pass.  A compiled pass may permute the qubits of its
// tile among the tile's bits for free at its store, so the low bits ("slots")
// are handed to qubits with pending work.  Invariant: slot i holds its own
// qubit i, or a visitor v while qubit i waits on v's own bit; every other qubit
// is on its own bit.  So at most 2 * low qubits are displaced, and one tile
// always suffices to restore the layout.
//
// After a pass, a lookahead picks the next pass's tile without the slot
// constraint; a slot whose holder the next tile does not want takes a wanted
// qubit v' of this tile that sits on its own bit (most pending work first):
//   slot i holds qubit i:          i -> bit v', v' -> slot i;
//   slot i holds v, i on bit v:    v' -> slot i, v -> bit v, i -> bit v'
//                                  (needs qubit i in this tile);
//   v' = i (qubit i comes back):   i -> slot i, v -> bit v.
// On the last pass the store restores every qubit it can reach to its own
// bit.  Records the store layout in p.store_phys.
void relabel_low(const Scheduler& sch, const std::vector<Node>& nodes, const std::vector<char>& in, uint32_t low,
                 uint32_t M, const std::vector<uint32_t>& all_q, std::vector<uint32_t>& phys,
                 std::vector<uint32_t>& at, Pass& p, bool last) {
    const uint32_t n = uint32_t(phys.size());
    std::vector<uint32_t> nphys(phys);
    std::vector<char> tile(n, 0);
    for (uint32_t b = 0; b < low; ++b) tile[b] = 1;
    for (uint32_t b : p.high) tile[b] = 1;
    std::vector<char> next(n, 0);
    if (!last) {
        if (!choose_set(sch, nodes, next, 0, M, all_q)) return;
        // When the next pass can finish the circuit, restore instead of
        // displacing more qubits, so that it needs no restore pass after it.
        std::vector<size_t> h = sch.head;
        sch.simulate(h, next, nullptr);
        bool all = true;
        for (uint32_t q = 0; q < n; ++q) all = all && h[q] == sch.queue[q].size();
        last = all;
    }
    if (last) {
        // Move every displaced qubit whose slot-chain lies in this tile home.
        for (uint32_t i = 0; i < low; ++i) {
            const uint32_t v = at[i];
            if (v != i && tile[phys[i]]) {
                nphys[v] = v;
                nphys[i] = i;
            }
        }
    } else {
        std::vector<char> taken(n, 0);
        for (uint32_t i = 0; i < low; ++i) {
            const uint32_t v = at[i];
            if (next[v]) continue;  // the holder is wanted: keep it
            if (v != i && next[i] && in[i]) {  // bring qubit i back
                nphys[i] = i;
                nphys[v] = v;
                continue;
            }
            if (v != i && !in[i]) continue;  // qubit i must move along: not reachable
            uint32_t best = n;
            size_t best_work = 0;
            for (uint32_t q = low; q < n; ++q) {
                if (!in[q] || !next[q] || taken[q] || phys[q] != q) continue;
                const size_t work = sch.queue[q].size() - sch.head[q];
                if (best == n || work > best_work) {
                    best = q;
                    best_work = work;
                }
            }
            if (best == n) continue;
            taken[best] = 1;
            nphys[best] = i;
            nphys[i] = best;  // qubit i waits on the visitor's own bit
            if (v != i) nphys[v] = v;
        }
    }
    bool same = true;
    for (uint32_t q = 0; q < n; ++q) same = same && nphys[q] == phys[q];
    if (same) return;
    // Store layout of the pass: tile bit -> index bit after the store.
    std::vector<uint32_t> tile_phys(M);
    for (uint32_t b = 0; b < low; ++b) tile_phys[b] = b;
    for (size_t k = 0; k < p.high.size(); ++k) tile_phys[low + k] = p.high[k];
    p.store_phys.resize(M);
    for (uint32_t t = 0; t < M; ++t) p.store_phys[t] = nphys[at[tile_phys[t]]];
    phys = nphys;
    for (uint32_t q = 0; q < n; ++q) at[phys[q]] = q;
}

// Appends a pass with no gates that moves every qubit back to its own bit
// (when the relabels left some displaced): its tile is the low bits, the
// displaced qubits' bits and, to make up M bits, further bits.
std::string restore_layout(uint32_t qubits, uint32_t low, uint32_t M, uint32_t RB, std::vector<uint32_t>& phys,
                           std::vector<uint32_t>& at, Plan& out) {
    std::vector<char> pin(qubits, 0);
    uint32_t count = low;
    for (uint32_t b = 0; b < low; ++b) pin[b] = 1;
    bool any = false;
    for (uint32_t q = 0; q < qubits; ++q)
        if (phys[q] != q) {
            any = true;
            for (uint32_t b : {phys[q], q})
                if (!pin[b]) {
                    pin[b] = 1;
                    ++count;
                }
        }
    if (!any) return "";
    if (count > M) return "layout restore needs more than one tile";
    for (uint32_t b = qubits; b-- > low && count < M;)
        if (!pin[b]) {
            pin[b] = 1;
            ++count;
        }
    Pass p{};
    p.kind = Pass::TILE;
    p.tile_bits = M;
    p.low = low;
    p.reg_bits = RB;
    for (uint32_t b = low; b < qubits; ++b)
        if (pin[b]) p.high.push_back(b);
    std::vector<uint32_t> tile_phys(M);
    for (uint32_t b = 0; b < low; ++b) tile_phys[b] = b;
    for (size_t k = 0; k < p.high.size(); ++k) tile_phys[low + k] = p.high[k];
    p.store_phys.resize(M);
    for (uint32_t t = 0; t < M; ++t) p.store_phys[t] = at[tile_phys[t]];  // each qubit to its own bit
    TileOp hdr;
    std::memset(&hdr, 0, sizeof(hdr));
    hdr.kind = TOP_PHASE;
    uint32_t packed = 0;
    for (uint32_t i = 0; i < RB; ++i) packed |= (low + i) << (4 * i);
    hdr.q0 = int32_t(packed);
    hdr.q1 = int32_t(RB);
    p.ops.push_back(hdr);
    ++out.stats.phases;
    out.passes.push_back(std::move(p));
    for (uint32_t q = 0; q < qubits; ++q) phys[q] = at[q] = q;
    return "";
}

}  // namespace

namespace {
std::string plan_passes(const std::vector<Node>& nodes, uint32_t qubits, uint32_t M, uint32_t low, uint32_t RB,
                        bool regk, const PlanOptions& opt, int tie, int block, Plan& out, int cap = 0,
                        int* max_cost = nullptr);
}  // namespace

uint32_t default_tile_bits(int precision) { return precision == 0 ? 11 : 12; }
uint32_t default_low_bits(int precision) { return precision == 0 ? 5 : 6; }

namespace {

// Plans by content: a circuit run again and again (a training loop, the
// gradient engine's per-call plans) is planned once.  The key covers every
// input of build_plan; a copied plan shares its compiled programs.
std::mutex g_plan_mu;
std::map<std::string, Plan> g_plans;
std::deque<std::string> g_plan_order;
constexpr size_t kPlanCacheEntries = 64;

// Two independent 64-bit hashes over a byte stream (FNV-1a and a
// multiply-xorshift), printed as 32 hex digits.
struct KeyHash {
    uint64_t h1 = 1469598103934665603ull, h2 = 0x9e3779b97f4a7c15ull;
    void mix(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h1 = (h1 ^ b[i]) * 1099511628211ull;
            h2 = (h2 + b[i] + 1) * 0xff51afd7ed558ccdull;
            h2 ^= h2 >> 31;
        }
    }
    void val(int64_t v) { mix(&v, sizeof v); }
    void dbl(double v) { mix(&v, sizeof v); }
    std::string hex() const {
        char b[40];
        std::snprintf(b, sizeof b, "%016llx%016llx", (unsigned long long)h1, (unsigned long long)h2);
        return b;
    }
};

void mix_gates(KeyHash& k, const std::vector<GateIn>& gates) {
    k.val(int64_t(gates.size()));
    for (const GateIn& g : gates) {
        k.val(g.kind);
        k.val(g.a);
        k.val(g.b);
        for (const cd& c : g.m) {
            k.dbl(c.re);
            k.dbl(c.im);
        }
        if (g.kind == 6)
            for (const cd& c : g.m4) {
                k.dbl(c.re);
                k.dbl(c.im);
            }
        k.val(g.axis);
        k.val(g.param);
        k.val(g.slot);
        k.val(g.flags);
    }
}

std::string plan_key(const std::vector<GateIn>& gates, uint32_t qubits, int precision, const PlanOptions& opt) {
    KeyHash k;
    k.val(qubits);
    k.val(precision);
    k.val(opt.fuse);
    k.val(opt.tile_qubits);
    k.val(opt.low_qubits);
    k.val(opt.max_block);
    k.val(opt.kernel);
    k.val(opt.reg_bits);
    k.val(opt.relabel);
    k.val(opt.compile);
    k.val(opt.balance);
    k.val(opt.scaled_rotations);
    k.val(relabel_enabled() ? 1 : 0);
    mix_gates(k, gates);
    return k.hex();
}
std::string build_plan_uncached(const std::vector<GateIn>& gates, uint32_t qubits, int precision,
                                const PlanOptions& opt, Plan& out);

// ---------------------------------------------------------------------------
// Dense-block plans (kernel = 4): each pass applies one k-qubit block matrix
// (2^k x 2^k, composed in FP64 on the host) on the matrix cores, after the
// diagonal gates that are ready at the start of the pass (elementwise
// factors).  Greedy over the gate DAG: a gate waits for the earlier gates on
// its qubits that do not commute with it (diagonal gates commute with each
// other); a pass first takes every ready diagonal gate, then starts its block
// at the first ready non-diagonal gate and takes, in circuit order, every
// further ready gate whose qubits fit in k.
// ---------------------------------------------------------------------------
inline cd c_mul(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cd c_add(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }

bool gate_is_diag(const GateIn& g) {
    if (g.kind == 2 || g.kind == 4) return true;
    return g.kind == 0 && g.m[1].re == 0.0 && g.m[1].im == 0.0 && g.m[2].re == 0.0 && g.m[2].im == 0.0;
}

// Diagonal of a diagonal gate as d[bit(b) * 2 + bit(a)] (a == b for one qubit).
void gate_diag(const GateIn& g, uint32_t* a, uint32_t* b, cd d[4]) {
    *a = g.a;
    if (g.kind == 0) {
        *b = g.a;
        d[0] = g.m[0];
        d[1] = d[2] = cd{0, 0};
        d[3] = g.m[3];
    } else if (g.kind == 2) {
        *b = g.b;
        d[0] = d[1] = d[2] = cd{1, 0};
        d[3] = cd{-1, 0};
    } else {
        *b = g.b;
        for (int k = 0; k < 4; ++k) d[k] = g.m[k];
    }
}

struct DenseBlock {
    std::vector<uint32_t> q;  // bit i of a block index = qubit q[i]
    std::vector<cd> U{cd{1, 0}};
    int pos(uint32_t qubit) const {
        for (size_t i = 0; i < q.size(); ++i)
            if (q[i] == qubit) return int(i);
        return -1;
    }
    void add_qubit(uint32_t qb) {
        const size_t d = size_t(1) << q.size(), d2 = 2 * d;
        std::vector<cd> V(d2 * d2, cd{0, 0});
        for (size_t r = 0; r < d2; ++r)
            for (size_t c = 0; c < d2; ++c)
                if ((r >> q.size()) == (c >> q.size())) V[r * d2 + c] = U[(r & (d - 1)) * d + (c & (d - 1))];
        U.swap(V);
        q.push_back(qb);
    }
    // U <- G U for a gate on block positions.
    void apply(const GateIn& g) {
        const size_t d = size_t(1) << q.size();
        const int pa = pos(g.a), pb = g.kind == 0 ? -1 : pos(g.b);
        const size_t ma = size_t(1) << pa, mb = pb >= 0 ? size_t(1) << pb : 0;
        std::vector<cd> v(d);
        for (size_t c = 0; c < d; ++c) {
            for (size_t r = 0; r < d; ++r) v[r] = U[r * d + c];
            switch (g.kind) {
                case 0:
                    for (size_t i = 0; i < d; ++i)
                        if (!(i & ma)) {
                            const cd x0 = v[i], x1 = v[i | ma];
                            v[i] = c_add(c_mul(g.m[0], x0), c_mul(g.m[1], x1));
                            v[i | ma] = c_add(c_mul(g.m[2], x0), c_mul(g.m[3], x1));
                        }
                    break;
                case 1:  // CNOT a -> b
                    for (size_t i = 0; i < d; ++i)
                        if ((i & ma) && !(i & mb)) std::swap(v[i], v[i | mb]);
                    break;
                case 2:
                    for (size_t i = 0; i < d; ++i)
                        if ((i & ma) && (i & mb)) v[i] = cd{-v[i].re, -v[i].im};
                    break;
                case 3:
                    for (size_t i = 0; i < d; ++i)
                        if ((i & ma) && !(i & mb)) std::swap(v[i], v[(i ^ ma) | mb]);
                    break;
                case 4:
                    for (size_t i = 0; i < d; ++i) v[i] = c_mul(g.m[((i & mb) ? 2 : 0) + ((i & ma) ? 1 : 0)], v[i]);
                    break;
                default: {  // 6: dense 4x4, basis index bit(b) * 2 + bit(a)
                    for (size_t i = 0; i < d; ++i)
                        if (!(i & ma) && !(i & mb)) {
                            const size_t idx[4] = {i, i | ma, i | mb, i | ma | mb};
                            cd x[4], y[4];
                            for (int k = 0; k < 4; ++k) x[k] = v[idx[k]];
                            for (int r = 0; r < 4; ++r) {
                                y[r] = cd{0, 0};
                                for (int k = 0; k < 4; ++k) y[r] = c_add(y[r], c_mul(g.m4[r * 4 + k], x[k]));
                            }
                            for (int k = 0; k < 4; ++k) v[idx[k]] = y[k];
                        }
                }
            }
            for (size_t r = 0; r < d; ++r) U[r * d + c] = v[r];
        }
    }
    // U <- U D for a diagonal D over block positions pa, pb (d[bit(pb) * 2 + bit(pa)]).
    void right_diag(int pa, int pb, const cd d[4]) {
        const size_t n = size_t(1) << q.size();
        for (size_t c = 0; c < n; ++c) {
            const cd f = d[(((c >> pb) & 1) << 1) | ((c >> pa) & 1)];
            for (size_t r = 0; r < n; ++r) U[r * n + c] = c_mul(U[r * n + c], f);
        }
    }
};

std::string plan_dense(const std::vector<GateIn>& gates, uint32_t qubits, uint32_t k, Plan& out) {
    if (k < 4 || k > 5) return "dense-block plans take max_block 4 or 5";
    if (qubits < k + 4) return "dense-block passes need at least max_block + 4 qubits";
    const size_t G = gates.size();
    std::vector<char> diag(G);
    for (size_t i = 0; i < G; ++i) {
        if (gates[i].kind == 5) return "dense-block plans take fixed gates only (no run-time rotations)";
        diag[i] = gate_is_diag(gates[i]) ? 1 : 0;
    }
    // Direct predecessors: per qubit, a non-diagonal gate waits for the run of
    // gates back to the last non-diagonal one; a diagonal gate for the last
    // non-diagonal gate only.
    std::vector<std::vector<size_t>> on(qubits), succ(G);
    std::vector<uint32_t> npred(G, 0);
    for (size_t i = 0; i < G; ++i) {
        const GateIn& g = gates[i];
        const uint32_t qs[2] = {g.a, g.b};
        const int nq = g.kind == 0 ? 1 : 2;
        std::vector<size_t> pred;
        for (int t = 0; t < nq; ++t) {
            const auto& L = on[qs[t]];
            for (size_t j = L.size(); j-- > 0;) {
                const size_t h = L[j];
                if (!diag[i] || !diag[h]) pred.push_back(h);
                if (!diag[h]) break;
            }
        }
        std::sort(pred.begin(), pred.end());
        pred.erase(std::unique(pred.begin(), pred.end()), pred.end());
        for (size_t h : pred) succ[h].push_back(i);
        npred[i] = uint32_t(pred.size());
        for (int t = 0; t < nq; ++t) on[qs[t]].push_back(i);
    }
    std::vector<char> done(G, 0);
    size_t left = G;
    auto finish = [&](size_t i) {
        done[i] = 1;
        --left;
        for (size_t s2 : succ[i]) --npred[s2];
    };
    while (left) {
        Pass p{};
        p.kind = Pass::DENSE;
        std::vector<size_t> pre;
        for (size_t i = 0; i < G; ++i)
            if (!done[i] && diag[i] && npred[i] == 0) {
                pre.push_back(i);
                finish(i);
            }
        DenseBlock B;
        size_t ops = pre.size();
        for (size_t i = 0; i < G; ++i) {
            if (done[i] || npred[i] != 0) continue;
            const GateIn& g = gates[i];
            const uint32_t qs[2] = {g.a, g.b};
            const int nq = g.kind == 0 ? 1 : 2;
            if (B.q.empty() && diag[i]) continue;  // a block starts at a non-diagonal gate
            int fresh = 0;
            for (int t = 0; t < nq; ++t) fresh += B.pos(qs[t]) < 0 ? 1 : 0;
            if (B.q.size() + fresh > k) continue;
            for (int t = 0; t < nq; ++t)
                if (B.pos(qs[t]) < 0) B.add_qubit(qs[t]);
            B.apply(g);
            finish(i);
            ++ops;
        }
        // Pad the block to k qubits with qubits it does not touch, highest first
        // (the low index bits stay column bits: contiguous columns).
        for (uint32_t qb = qubits; qb-- > 0 && B.q.size() < k;)
            if (B.pos(qb) < 0) B.add_qubit(qb);
        // Diagonal gates ready at the start of the pass: inside the block into
        // U (applied before it), otherwise per column or per block bit.
        for (size_t i : pre) {
            uint32_t a, b;
            cd d[4];
            gate_diag(gates[i], &a, &b, d);
            const int pa = B.pos(a), pb = B.pos(b);
            if (pa >= 0 && pb >= 0) {
                B.right_diag(pa, pb, d);
            } else if (pa >= 0 || pb >= 0) {
                DiagPre e{};
                if (pa >= 0) {
                    e.a = uint32_t(pa);
                    e.b = b;
                    for (int t = 0; t < 4; ++t) e.d[t] = d[t];
                } else {  // swap the roles: block position first
                    e.a = uint32_t(pb);
                    e.b = a;
                    e.d[0] = d[0];
                    e.d[1] = d[2];
                    e.d[2] = d[1];
                    e.d[3] = d[3];
                }
                p.pre_blk.push_back(e);
            } else {
                DiagPre e{a, b, {d[0], d[1], d[2], d[3]}};
                p.pre_col.push_back(e);
            }
        }
        // Block qubits in ascending order.
        std::vector<uint32_t> order(B.q.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = uint32_t(i);
        std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return B.q[x] < B.q[y]; });
        std::vector<uint32_t> newpos(B.q.size());
        for (size_t i = 0; i < order.size(); ++i) {
            p.block.push_back(B.q[order[i]]);
            newpos[order[i]] = uint32_t(i);
        }
        const size_t d = size_t(1) << k;
        auto remap = [&](size_t x) {
            size_t y = 0;
            for (size_t i = 0; i < k; ++i)
                if ((x >> i) & 1) y |= size_t(1) << newpos[i];
            return y;
        };
        p.U.assign(d * d, cd{0, 0});
        for (size_t r = 0; r < d; ++r)
            for (size_t c = 0; c < d; ++c) p.U[remap(r) * d + remap(c)] = B.U[r * d + c];
        for (DiagPre& e : p.pre_blk) e.a = newpos[e.a];
        out.stats.max_ops_per_pass = std::max<uint64_t>(out.stats.max_ops_per_pass, ops);
        out.passes.push_back(std::move(p));
        ++out.stats.dense_blocks;
    }
    out.stats.passes = out.passes.size();
    out.stats.ops_applied = G;
    return "";
}

}  // namespace

std::string circuit_key(const std::vector<GateIn>& gates, uint32_t qubits) {
    KeyHash k;
    k.val(qubits);
    mix_gates(k, gates);
    return k.hex();
}

std::string build_plan(const std::vector<GateIn>& gates, uint32_t qubits, int precision, const PlanOptions& opt,
                       Plan& out) {
    const std::string key = plan_key(gates, qubits, precision, opt);
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = g_plans.find(key);
        if (it != g_plans.end()) {
            out = it->second;
            return "";
        }
    }
    std::string e = build_plan_uncached(gates, qubits, precision, opt, out);
    if (!e.empty()) return e;
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (g_plans.emplace(key, out).second) {
        g_plan_order.push_back(key);
        if (g_plan_order.size() > kPlanCacheEntries) {
            g_plans.erase(g_plan_order.front());
            g_plan_order.pop_front();
        }
    }
    return "";
}

namespace {
std::string build_plan_uncached(const std::vector<GateIn>& gates, uint32_t qubits, int precision,
                                const PlanOptions& opt, Plan& out) {
    out = Plan{};
    out.qubits = qubits;
    out.precision = precision;
    out.stats.ops_in = gates.size();
    out.circuit_key = circuit_key(gates, qubits);
    for (const GateIn& g : gates) out.runtime_rotations += g.kind == 5 ? 1 : 0;
    out.origin = out.runtime_rotations ? PLAN_PARAMETERIZED : PLAN_CIRCUIT;
    if (qubits == 0) return "a state vector must contain at least one qubit";
    for (const GateIn& g : gates) {
        const bool two = g.kind != 0 && g.kind != 5;
        if (g.a >= qubits || (two && g.b >= qubits)) return "qubit out of range";
        if (two && g.a == g.b) return "two-qubit operation requires distinct qubits";
    }
    if (opt.fuse && opt.kernel == 4) return plan_dense(gates, qubits, opt.max_block ? opt.max_block : 5, out);
    if (!opt.fuse) {
        for (const GateIn& g : gates)
            if (g.kind == 5 || g.kind == 6) return "per-instance rotations and dense two-qubit ops need the tile lowering";
        for (const GateIn& g : gates) {
            Pass p{};
            if (g.kind == 0) {
                p.kind = Pass::SWEEP_1Q;
                p.q0 = g.a;
                std::memcpy(p.m, g.m, sizeof(p.m));
            } else if (g.kind == 4) {
                p.kind = Pass::SWEEP_DIAG;
                p.q0 = g.a;
                p.q1 = g.b;
                std::memcpy(p.m, g.m, sizeof(p.m));
            } else {
                p.kind = Pass::SWEEP_2Q;
                p.op2 = g.kind == 1 ? 0 : (g.kind == 2 ? 1 : 2);
                p.q0 = g.a;
                p.q1 = g.b;
            }
            out.passes.push_back(p);
        }
        out.stats.passes = out.passes.size();
        out.stats.ops_applied = gates.size();
        out.stats.max_ops_per_pass = gates.empty() ? 0 : 1;
        return "";
    }

    const std::vector<Node> nodes = lower(gates, qubits);
    out.stats.ops_applied = nodes.size();
    uint32_t M = opt.tile_qubits ? opt.tile_qubits : default_tile_bits(precision);
    uint32_t low = opt.low_qubits ? opt.low_qubits : default_low_bits(precision);
    if (M > 14) return "tile_qubits must be at most 14";
    // Register kernel: 2^RB amplitudes per thread, 2^(M-RB) <= 256 threads.
    // Compiled passes take the same plan with 2 <= RB <= 5 and up to 1024 threads.
    uint32_t RB = opt.reg_bits ? opt.reg_bits : (precision == 0 ? 3 : 4);
    const bool want_compiled = opt.kernel == 3 || (opt.kernel == 0 && (qubits >= kCompileMinQubits || opt.compile));
    bool regk = opt.kernel != 1 && M <= qubits && RB >= 3 && RB <= 4 && M >= RB + 6 && M - RB <= 8;
    if (want_compiled) regk = M <= qubits && RB >= 2 && RB <= 5 && M >= RB + 6 && M - RB <= 10;
    if (opt.kernel == 2 && !regk) return "register kernel needs tile_qubits <= qubits and 6 <= tile - reg_bits <= 8";
    if (opt.kernel == 3 && !regk)
        return "compiled passes need tile_qubits <= qubits, 2 <= reg_bits <= 5 and 6 <= tile - reg_bits <= 10";
    out.compiled = regk && want_compiled ? (opt.kernel == 3 ? 2 : 1) : 0;
    // Compiled passes need only one 128-byte run of contiguous amplitudes
    // per tile for full cache lines; the freed tile slots take gate qubits.
    // Compiled tiles are one qubit larger than the generic kernels' (12
    // qubits complex128, 13 complex64: 64 KiB of LDS, 512-thread workgroups
    // with 3 or 4 register bits): fewer passes at the same streaming rate.
    if (out.compiled && !opt.low_qubits) low = precision == 0 ? 3 : 4;
    if (out.compiled && !opt.tile_qubits && qubits >= M + 1 && M + 1 - RB <= 10) ++M;
    if (M > qubits) M = qubits;
    if (low > M) low = M;
    // A tile smaller than the state must leave room for the two qubits of a
    // gate outside the contiguous low block.
    if (M < qubits && low + 2 > M) low = M >= 2 ? M - 2 : 0;

    // Compiled plans are built under several tie-break policies of the set
    // choice and the one with the fewest passes (then phases) is kept:
    // planning takes milliseconds, a pass over a large state takes as long.
    const Plan base = out;
    std::vector<std::pair<int, int>> policies = {{0, 1}};
    if (out.compiled) policies = {{0, 1}, {1, 1}, {2, int(qubits)}, {2, int(qubits + qubits / 3)}};
    std::string first_err;
    bool have = false;
    Plan best;
    int best_max = 0;  // node_cost of the chosen plan's heaviest pass
    for (const auto& pol : policies) {
        Plan trial = base;
        int mx = 0;
        std::string e = plan_passes(nodes, qubits, M, low, RB, regk, opt, pol.first, pol.second, trial, 0, &mx);
        if (!e.empty()) {
            if (first_err.empty()) first_err = e;
            continue;
        }
        if (!have || trial.passes.size() < best.passes.size() ||
            (trial.passes.size() == best.passes.size() && trial.stats.phases < best.stats.phases)) {
            best = std::move(trial);
            best_max = mx;
            have = true;
        }
    }
    if (!have) return first_err;
    if (out.compiled && opt.balance && best.passes.size() >= 2) {
        // Caps at a share of the arithmetic per pass: a capped plan with no
        // more passes and a lighter heaviest pass replaces the greedy one.
        int total = 0;
        for (const Node& n : nodes) total += node_cost(n);
        const size_t P = best.passes.size();
        for (int pct : {120, 140, 160}) {
            const int cap = int((int64_t(total) * pct + int64_t(P) * 100 - 1) / (int64_t(P) * 100));
            for (const auto& pol : policies) {
                Plan trial = base;
                int mx = 0;
                if (!plan_passes(nodes, qubits, M, low, RB, regk, opt, pol.first, pol.second, trial, cap, &mx).empty())
                    continue;
                if (trial.passes.size() <= P && mx < best_max) {
                    best = std::move(trial);
                    best_max = mx;
                }
            }
        }
    }
    out = std::move(best);
    return "";
}

}  // namespace

namespace {

// The pass loop of build_plan under one tie-break policy of the set choice.
std::string plan_passes(const std::vector<Node>& nodes, uint32_t qubits, uint32_t M, uint32_t low, uint32_t RB,
                        bool regk, const PlanOptions& opt, int tie, int block, Plan& out, int cap, int* max_cost) {
    Scheduler sch(nodes, qubits, regk ? 1 : 0);
    sch.tie = tie;
    sch.block = block > 0 ? block : 1;
    sch.cap = cap;
    if (max_cost) *max_cost = 0;
    std::vector<uint32_t> all_q(qubits);
    for (uint32_t q = 0; q < qubits; ++q) all_q[q] = q;
    // Layout: phys[q] = the index bit that holds logical qubit q.  Compiled
    // passes may relabel (see relabel_low below); other plans keep the
    // identity.
    const bool relabel = out.compiled && opt.relabel && low > 0 && M < qubits && relabel_enabled();
    std::vector<uint32_t> phys(qubits), at(qubits);
    for (uint32_t q = 0; q < qubits; ++q) phys[q] = at[q] = q;
    size_t done = 0;
    while (done < nodes.size()) {
        std::vector<char> in(qubits, 0);
        for (uint32_t b = 0; b < low; ++b) in[at[b]] = 1;
        bool chosen = false;
        if (relabel) {
            // When the rest of the circuit fits one pass that also holds the
            // displaced qubits, that pass restores the layout at its store
            // and no restore pass is needed.
            std::vector<char> fin(in);
            uint32_t fsize = low;
            for (uint32_t q = 0; q < qubits; ++q)
                if (phys[q] != q && !fin[q]) {
                    fin[q] = 1;
                    ++fsize;
                }
            if (fsize > low && fsize <= M && choose_set(sch, nodes, fin, fsize, M, all_q)) {
                std::vector<size_t> h = sch.head;
                sch.simulate(h, fin, nullptr);
                bool all = true;
                for (uint32_t q = 0; q < qubits; ++q) all = all && h[q] == sch.queue[q].size();
                if (all) {
                    in = fin;
                    chosen = true;
                }
            }
        }
        if (!chosen && !(out.compiled ? choose_set_lookahead(sch, nodes, in, low, M, all_q, at)
                                      : choose_set(sch, nodes, in, low, M, all_q)))
            return "planner made no progress";
        std::vector<int> order;
        sch.simulate(sch.head, in, &order);
        if (order.empty()) return "planner made no progress";
        done += order.size();
        if (max_cost) {
            int c = 0;
            for (int id : order) c += node_cost(nodes[size_t(id)]);
            *max_cost = std::max(*max_cost, c);
        }
        Pass p{};
        p.kind = Pass::TILE;
        p.tile_bits = M;
        p.low = low;
        // The pass in index-bit terms: tile bits low..M-1 are the held bits of
        // the chosen qubits in increasing order.
        std::vector<char> pin(qubits, 0);
        for (uint32_t q = 0; q < qubits; ++q)
            if (in[q]) pin[phys[q]] = 1;
        std::vector<int> tile_bit(qubits, -1);
        for (uint32_t b = 0; b < low; ++b) tile_bit[b] = int(b);
        uint32_t k = low;
        for (uint32_t b = low; b < qubits; ++b)
            if (pin[b]) {
                p.high.push_back(b);
                tile_bit[b] = int(k++);
            }
        std::vector<Node> pn;
        pn.reserve(order.size());
        for (int id : order) {
            Node x = nodes[size_t(id)];
            for (int j = 0; j < x.nq; ++j) x.q[j] = phys[x.q[j]];
            pn.push_back(x);
        }
        std::vector<int> porder(pn.size());
        for (size_t i = 0; i < pn.size(); ++i) porder[i] = int(i);
        if (!regk) {
            for (const Node& x : pn) p.ops.push_back(make_tile_op(x, tile_bit));
        } else {
            p.reg_bits = RB;
            std::string e = build_phases(pn, porder, qubits, tile_bit, RB, p.ops, out.stats.phases);
            if (!e.empty()) return e;
        }
        if (relabel) relabel_low(sch, nodes, in, low, M, all_q, phys, at, p, done == nodes.size());
        out.stats.max_ops_per_pass = std::max<uint64_t>(out.stats.max_ops_per_pass, order.size());
        out.passes.push_back(std::move(p));
    }
    if (relabel) {
        std::string e = restore_layout(qubits, low, M, RB, phys, at, out);
        if (!e.empty()) return e;
    }
    for (const Pass& p : out.passes)
        if (!p.store_phys.empty()) ++out.relabeled;
    if (opt.scaled_rotations && out.compiled && out.runtime_rotations) {
        bool all = true;
        for (const Pass& p : out.passes)
            for (const TileOp& o : p.ops)
                if (o.kind == TOP_ROT && o.q1 != 2 && (p.kind != Pass::TILE || p.reg_bits == 0)) all = false;
        out.scaled_rotations = all ? 1 : 0;
    }
    out.stats.passes = out.passes.size();
    return "";
}

}  // namespace

}  // namespace qf
