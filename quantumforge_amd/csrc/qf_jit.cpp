// Pass compiler: source generation for compiled tile passes, hiprtc
// compilation and the in-process / on-disk caches (interface in qf_jit.h).
#include "qf_jit.h"

#include <dirent.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <thread>

namespace qf {
namespace {

std::string hexlit(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%a", v);
    return std::string("R(") + b + ")";
}

// One product coefficient * variable of a linear combination.  coef is a
// source expression, or empty for 1.
struct Term {
    std::string coef;
    bool neg;
    std::string var;
};

// A zero coefficient yields a term with no variable, which lincomb drops
// (x * 0 cannot be folded by the compiler: x may be infinite or NaN).
Term lit_term(double c, const std::string& var) {
    if (c == 0.0) return Term{std::string(), false, std::string()};
    const double a = std::fabs(c);
    return Term{a == 1.0 ? std::string() : hexlit(a), std::signbit(c), var};
}

std::string lincomb(const std::vector<Term>& ts) {
    std::string s;
    for (const Term& t : ts) {
        if (t.var.empty()) continue;
        const std::string body = t.coef.empty() ? t.var : t.coef + " * " + t.var;
        if (s.empty())
            s = t.neg ? "-" + body : body;
        else
            s += (t.neg ? " - " : " + ") + body;
    }
    return s.empty() ? "R(0)" : s;
}

struct Bit {
    enum Kind { REG, THREAD, BLOCK } kind;
    int idx;  // register index, thread bit, or global qubit
};

std::string hex64(unsigned long long v) {
    char b[24];
    std::snprintf(b, sizeof b, "%016llx", v);
    return b;
}

class PassGen {
    using zc = std::complex<double>;

public:
    // ns = 2: the adjoint form, which carries a second state (the costate
    // lambda) through the same ops and records Im <lambda| G |psi> for the
    // TF_RECORD rotations before applying them.
    PassGen(const Pass& p, uint32_t n, int precision, const std::string& name, int ns = 1, bool scaled = false)
        : p_(p), n_(n), dbl_(precision == 0), name_(name), ns_(ns), scaled_(scaled) {}

    bool run(std::string* err) {
        if (p_.kind != Pass::TILE || p_.reg_bits == 0 || p_.ops.empty() || p_.ops[0].kind != TOP_PHASE)
            return fail(err, "not a register-kernel tile pass");
        M_ = p_.tile_bits;
        RB_ = p_.reg_bits;
        if (RB_ > 6 || M_ < RB_ || M_ - RB_ > 10 || M_ > 16 || M_ > n_) return fail(err, "unsupported tile shape");
        NR_ = 1u << RB_;
        NT_ = 1u << (M_ - RB_);
        LOGB_ = n_ - M_;
        tile_of_q_.assign(n_, -1);
        q_of_tile_.assign(M_, 0);
        for (uint32_t b = 0; b < p_.low; ++b) {
            tile_of_q_[b] = int(b);
            q_of_tile_[b] = b;
        }
        for (size_t k = 0; k < p_.high.size(); ++k) {
            tile_of_q_[p_.high[k]] = int(p_.low + k);
            q_of_tile_[p_.low + k] = p_.high[k];
        }
        bool any_rot = false;
        size_t phases = 0;
        std::vector<int> rec_slots;
        for (const TileOp& op : p_.ops) {
            if (op.kind == TOP_ROT || op.kind == TOP_ROTZZ) any_rot = true;
            if (op.kind == TOP_PHASE) ++phases;
            if (ns_ == 2 && (op.kind == TOP_ROT || op.kind == TOP_ROTZZ) && (op.flags & TF_RECORD))
                rec_slots.push_back(op.slot);
        }
        const uint32_t W = NT_ >= 64 ? NT_ / 64 : 1;
        const char* R = dbl_ ? "double" : "float";
        o_ << "extern \"C\" __global__ __launch_bounds__(" << NT_ << "@LB@) void " << name_
           << "(C* __restrict__ s, const double* __restrict__ cs, unsigned nslots"
           << (ns_ == 2 ? ", C* __restrict__ lam, double* __restrict__ partial" : "") << ") {\n";
        relabel_ = !p_.store_phys.empty();
        if (relabel_ && !check_store_layout()) return fail(err, "store layout is not a permutation of the tile bits");
        plan_half_exchanges();
        if (phases > 1 || relabel_) o_ << "  __shared__ C xb[" << ns_ * (1u << (half_ ? M_ - 1 : M_)) << "];\n";
        if (!rec_slots.empty()) o_ << "  __shared__ double qred[" << rec_slots.size() * W << "];\n";
        o_ << "  typedef " << R << " R;\n";
        o_ << "  const unsigned t = threadIdx.x;\n";
        // Workgroup x of a row runs on XCD x mod 8 (round-robin dispatch);
        // each XCD takes a contiguous run of the row's tiles, so tiles that
        // differ in the lowest block bits (the neighbouring runs in memory)
        // are fetched by one XCD close together in time.
        o_ << "  const unsigned xg = gridDim.x >= 8u && (gridDim.x & 7u) == 0u ? gridDim.x >> 3 : 0u;\n";
        o_ << "  const unsigned long long b = (unsigned long long)(xg ? (blockIdx.x & 7u) * xg + (blockIdx.x >> 3) : "
              "blockIdx.x) + (unsigned long long)blockIdx.y * gridDim.x;\n";
        o_ << "  const unsigned long long inst = b >> " << LOGB_ << ";\n";
        o_ << "  const unsigned long long bl = b & " << mask64(LOGB_) << "ull;\n";
        o_ << "  const unsigned long long base = (inst << " << n_ << ")" << block_scatter() << ";\n";
        if (any_rot) o_ << "  const double* __restrict__ csr = cs + inst * 2ull * nslots;\n";
        else o_ << "  (void)cs; (void)nslots;\n";
        for (uint32_t k = 0; k < NR_; ++k) o_ << "  R r" << k << ", i" << k << ";\n";
        if (ns_ == 2)
            for (uint32_t k = 0; k < NR_; ++k) o_ << "  R lr" << k << ", li" << k << ";\n";
        acc_.assign(1 + 2 * NR_, Acc{});
        acc_[0].r = "ar";
        acc_[0].i = "ai";
        o_ << "  R ar = R(1), ai = R(0);\n";
        // Phase (in 2^-64 turns) of the literal diagonal factors that depend
        // on tile-index bits only: the same for the whole tile, so it is
        // added up in scalar integer registers and applied once at the end.
        o_ << "  unsigned long long uph = 0ull;\n";
        for (uint32_t b = 0; b < RB_; ++b)
            for (int v = 0; v < 2; ++v) {
                Acc& a = acc_[1 + 2 * b + v];
                a.r = "a" + std::to_string(b) + std::to_string(v) + "r";
                a.i = "a" + std::to_string(b) + std::to_string(v) + "i";
                o_ << "  R " << a.r << " = R(1), " << a.i << " = R(0);\n";
            }
        var_.resize(NR_);
        for (uint32_t k = 0; k < NR_; ++k) var_[k] = int(k);
        set_phase(p_.ops[0]);
        // Load with the first phase's mapping.
        for (int st = 0; st < ns_; ++st) {
            o_ << "  {\n    C* __restrict__ sp = " << (st ? "lam" : "s") << " + base + " << thread_global() << ";\n";
            for (uint32_t k = 0; k < NR_; ++k)
                o_ << "    { const C v = sp[" << reg_global(k) << "ull]; " << reg_name(st, 'r', k) << " = v.re; "
                   << reg_name(st, 'i', k) << " = v.im; }\n";
            o_ << "  }\n";
        }
        if (scaled_) scaled_flags();
        bool lds_used = false;
        uint32_t rec_index = 0;
        for (size_t i = 1; i < p_.ops.size(); ++i) {
            const TileOp& op = p_.ops[i];
            cur_op_ = i;
            ++opn_;
            switch (op.kind) {
                case TOP_PHASE: exchange(layout_of(op), lds_used); lds_used = true; break;
                case TOP_MAT1: mat1(op); break;
                case TOP_DIAG: diag(op); break;
                case TOP_CNOT: cnot(op); break;
                case TOP_SWAP: swap(op); break;
                case TOP_MAT2: mat2(op); break;
                case TOP_ROT:
                    if (ns_ == 2 && (op.flags & TF_RECORD)) record(op, rec_index++, W);
                    rot(op);
                    break;
                case TOP_ROTZZ:
                    if (ns_ == 2 && (op.flags & TF_RECORD)) record(op, rec_index++, W);
                    rotzz(op);
                    break;
                default: return fail(err, "op kind " + std::to_string(op.kind) + " is not compiled");
            }
        }
        if (g_ != zc(1.0, 0.0)) acc_mul_lit(acc_[0], g_.real(), g_.imag());
        g_ = 1.0;
        if (uph_) {
            o_ << "  { const double th = double((long long)uph) * 0x1.921fb54442d18p-62;  // 2 pi 2^-64\n"
                  "    const R upr = R(cos(th)), upi = R(sin(th));\n";
            acc_mul(acc_[0], "upr", "upi");
            o_ << "  }\n";
        }
        acc_flush_all();
        // A relabeling pass first moves the tile into the store layout, whose
        // lanes 0..low-1 carry the tile bits stored to index bits 0..low-1.
        if (relabel_) exchange(store_layout(), lds_used);
        const std::vector<uint32_t>& pos = relabel_ ? p_.store_phys : q_of_tile_;
        for (cur_ = 0; cur_ < ns_; ++cur_) {
            o_ << "  {\n    C* __restrict__ sp = " << (cur_ ? "lam" : "s") << " + base + " << thread_scatter(true, pos)
               << ";\n";
            for (uint32_t k = 0; k < NR_; ++k)
                o_ << "    sp[" << reg_global(k, pos) << "ull] = C{" << rv(k) << ", " << iv(k) << "};\n";
            o_ << "  }\n";
        }
        cur_ = 0;
        flush_records();
        if (!rec_slots.empty()) {
            // The tile's overlap for each recorded rotation: its waves' sums
            // added in wave order, into the block's row of the partial table.
            o_ << "  __syncthreads();\n  for (unsigned r = t; r < " << rec_slots.size() << "u; r += " << NT_
               << "u) {\n    double q = 0.0;\n";
            o_ << "    for (unsigned w = 0; w < " << W << "u; ++w) q += qred[r * " << W << "u + w];\n";
            std::string sl = "0u";
            for (size_t r = rec_slots.size(); r-- > 0;)
                sl = r + 1 == rec_slots.size() ? std::to_string(rec_slots[r]) + "u"
                                               : "(r == " + std::to_string(r) + "u ? " + std::to_string(rec_slots[r]) +
                                                     "u : " + sl + ")";
            o_ << "    partial[b * nslots + " << sl << "] = q;\n  }\n";
        }
        o_ << "}\n";
        src = o_.str();
        // 512-thread workgroups whose shared arrays let two of them share a
        // CU (160 KiB of LDS) ask for four waves per SIMD, i.e. at most 128
        // VGPRs: the two-state passes otherwise land just above and halve
        // their resident waves.
        const size_t lds = size_t(phases > 1 || relabel_ ? ns_ * (1u << (half_ ? M_ - 1 : M_)) : 0) *
                               (dbl_ ? 16 : 8) +
                           rec_slots.size() * W * 8;
        const std::string lb = (NT_ == 512 && 2 * lds <= 160 * 1024) ? ", 4" : "";
        src.replace(src.find("@LB@"), 4, lb);
        return true;
    }

    std::string src;

private:
    // A phase layout: register index of each tile bit (or -1) and thread bit
    // of each tile bit (or -1).
    struct Layout {
        std::vector<int> reg, thr;
    };

    const Pass& p_;
    uint32_t n_;
    bool dbl_;
    std::string name_;
    int ns_ = 1;                  // states carried (2: psi and the costate lambda)
    bool scaled_ = false;         // X/Y rotation rows hold (scale, ratio)
    // Scaled rotations of the pass: op index -> position in pass order (its
    // form flag is bit pos % 32 of word fw<pos / 32>), and the variables of
    // the ratios already loaded.
    std::map<size_t, uint32_t> sc_pos_;
    std::vector<size_t> sc_ops_;
    std::map<size_t, std::string> sc_u_;
    size_t cur_op_ = 0;
    int cur_ = 0;                 // the state rv()/iv() name
    uint32_t M_ = 0, RB_ = 0, NR_ = 0, NT_ = 0, LOGB_ = 0;
    std::vector<int> tile_of_q_;
    std::vector<uint32_t> q_of_tile_;
    int rb_[8] = {0};
    std::vector<int> reg_of_tile_, thr_of_tile_;
    std::vector<int> var_;
    zc g_{1.0, 0.0};              // the true amplitudes are g_ times the tile's (take_scale)
    bool relabel_ = false;        // the store permutes the tile's bits (Pass::store_phys)
    bool uph_ = false;            // some diagonal factor went into the tile phase uph
    bool half_ = false;           // exchanges through a half-tile buffer
    // Half-buffer exchanges in emission order: the round bit, and for an
    // exchange whose two layouts share no thread bit, an intermediate layout
    // (and its round bit) it passes through.
    struct HalfX {
        int c = -1;
        bool hop = false;
        Layout mid;
        int c_mid = -1;
    };
    std::vector<HalfX> half_x_;
    size_t xi_ = 0;               // exchanges emitted so far
    std::ostringstream o_;
    int opn_ = 0;
    // Diagonal factor accumulators.  acc_[0]: the product of the diagonal
    // factors that depend on no register bit: it scales all of a thread's
    // amplitudes alike, so it commutes with every op of the phase and is
    // applied once, before the next exchange or the store.  acc_[1 + 2i + v]:
    // factors that depend on register bit i (value v) and on thread or block
    // bits: they commute with everything but ops that are not diagonal on
    // bit i, and are applied before those.
    struct Acc {
        std::string r, i;
        bool on = false, real = true;
    };
    std::vector<Acc> acc_;
    // Recorded overlaps waiting for their wave sums: a thread's share of
    // records rec0_, rec0_ + 1, ... (variable names), summed over the wave
    // together once kRecBatch of them are pending.
    static constexpr size_t kRecBatch = 4;
    std::vector<std::string> pend_;
    uint32_t rec0_ = 0;
    uint32_t recw_ = 1;  // waves per workgroup (rows of the shared table)

    static bool fail(std::string* err, const std::string& m) {
        if (err) *err = m;
        return false;
    }
    static std::string mask64(uint32_t bits) {
        char b[32];
        std::snprintf(b, sizeof b, "0x%llxu", bits >= 64 ? ~0ull : ((1ull << bits) - 1));
        std::string s(b);
        s.pop_back();  // the caller appends the suffix
        return s;
    }
    // Register variables: state 0 r<k>/i<k>, state 1 (lambda) lr<k>/li<k>.
    static std::string reg_name(int st, char part, uint32_t k) {
        return std::string(st ? "l" : "") + part + std::to_string(k);
    }
    std::string rv(uint32_t k) const { return reg_name(cur_, 'r', uint32_t(var_[k])); }
    std::string iv(uint32_t k) const { return reg_name(cur_, 'i', uint32_t(var_[k])); }

    // Bits of the block index go, in order, to the qubits outside the tile.
    std::string block_scatter() const {
        std::string s;
        uint32_t bit = 0;
        uint32_t g = 0;
        while (g < n_) {
            if (tile_of_q_[g] >= 0) {
                ++g;
                continue;
            }
            uint32_t len = 0;
            while (g + len < n_ && tile_of_q_[g + len] < 0) ++len;
            s += " | ((bl >> " + std::to_string(bit) + ") & " + mask64(len) + "ull) << " + std::to_string(g);
            bit += len;
            g += len;
        }
        return s;
    }

    Layout layout_of(const TileOp& hdr) const {
        Layout L;
        L.reg.assign(M_, -1);
        for (uint32_t i = 0; i < RB_; ++i) L.reg[(uint32_t(hdr.q0) >> (4 * i)) & 15u] = int(i);
        L.thr = assign_threads(L.reg);
        return L;
    }
    void set_layout(const Layout& L) {
        reg_of_tile_ = L.reg;
        thr_of_tile_ = L.thr;
        for (uint32_t tb = 0; tb < M_; ++tb)
            if (L.reg[tb] >= 0) rb_[L.reg[tb]] = int(tb);
    }
    void set_phase(const TileOp& hdr) { set_layout(layout_of(hdr)); }

    // The store layout of a relabeling pass: thread bits 0..low-1 on the tile
    // bits stored to index bits 0..low-1 (so each run of 2^low lanes writes
    // one contiguous run), the register and thread bits of the other tile
    // bits those of the last phase where possible.
    bool check_store_layout() const {
        if (p_.store_phys.size() != M_) return false;
        std::vector<uint32_t> a(p_.store_phys), b(q_of_tile_);
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        if (a != b) return false;
        for (uint32_t i = 0; i < p_.low; ++i)
            if (std::find(p_.store_phys.begin(), p_.store_phys.end(), i) == p_.store_phys.end()) return false;
        return true;
    }
    Layout store_layout() const {
        const uint32_t low = p_.low;
        std::vector<int> dest(low, -1);
        std::vector<char> is_dest(M_, 0);
        for (uint32_t t = 0; t < M_; ++t)
            if (p_.store_phys[t] < low) {
                dest[p_.store_phys[t]] = int(t);
                is_dest[t] = 1;
            }
        const TileOp* last = nullptr;
        for (const TileOp& op : p_.ops)
            if (op.kind == TOP_PHASE) last = &op;
        Layout L;
        L.reg.assign(M_, -1);
        uint32_t nreg = 0;
        const Layout prev = layout_of(*last);
        for (uint32_t i = 0; i < RB_; ++i)
            for (uint32_t t = 0; t < M_; ++t)
                if (prev.reg[t] == int(i) && !is_dest[t]) L.reg[t] = int(nreg++);
        // A register bit given up to the store's lanes is replaced by a tile
        // bit whose thread bit the lanes take anyway (else the highest free).
        for (uint32_t t = 0; t < M_ && nreg < RB_; ++t)
            if (L.reg[t] < 0 && !is_dest[t] && prev.thr[t] >= 0 && prev.thr[t] < int(low)) L.reg[t] = int(nreg++);
        for (uint32_t t = M_; t-- > 0 && nreg < RB_;)
            if (L.reg[t] < 0 && !is_dest[t]) L.reg[t] = int(nreg++);
        const int T = int(M_ - RB_);
        L.thr.assign(M_, -1);
        std::vector<char> taken(T, 0);
        for (uint32_t i = 0; i < low; ++i) {
            L.thr[dest[i]] = int(i);
            taken[i] = 1;
        }
        // Every other tile bit keeps its thread bit from the last phase where
        // it can, so the store exchange finds a common bit for the half buffer.
        for (uint32_t t = 0; t < M_; ++t)
            if (L.reg[t] < 0 && L.thr[t] < 0 && prev.thr[t] >= int(low) && !taken[prev.thr[t]]) {
                L.thr[t] = prev.thr[t];
                taken[prev.thr[t]] = 1;
            }
        int u = 0;
        for (uint32_t t = 0; t < M_; ++t) {
            if (L.reg[t] >= 0 || L.thr[t] >= 0) continue;
            while (taken[u]) ++u;
            L.thr[t] = u++;
        }
        return L;
    }

    // Thread bits of a phase: the top tile positions M-1, M-2, M-3 keep the
    // wave-level thread bits T-1, T-2, T-3 (T = M - RB thread bits, those
    // above 6) whenever they are not register bits, so that an exchange finds
    // a wave bit with the same meaning before and after it; the other tile
    // positions take the remaining thread bits in increasing order, so the
    // lane bits stay on the lowest (contiguous) positions.
    std::vector<int> assign_threads(const std::vector<int>& reg) const {
        const int T = int(M_ - RB_);
        const int nres = std::max(0, std::min(3, T - 6));
        std::vector<int> thr(M_, -1);
        std::vector<char> taken(T, 0);
        for (int i = 0; i < nres; ++i) {
            const uint32_t pos = M_ - 1 - uint32_t(i);
            if (reg[pos] < 0) {
                thr[pos] = T - 1 - i;
                taken[T - 1 - i] = 1;
            }
        }
        int u = 0;
        for (uint32_t tb = 0; tb < M_; ++tb) {
            if (reg[tb] >= 0 || thr[tb] >= 0) continue;
            while (taken[u]) ++u;
            thr[tb] = u++;
        }
        return thr;
    }

    // Thread t's tile index from the phase's thread-bit map (assign_threads).
    // Runs of consecutive thread bits landing on consecutive positions are one
    // shift-and-mask each.  `global` selects global qubit positions instead of
    // tile bit positions.
    std::string thread_scatter(bool global) const { return thread_scatter(global, q_of_tile_); }
    std::string thread_scatter(bool global, const std::vector<uint32_t>& qpos) const {
        std::vector<std::pair<int, uint32_t>> pos;  // (thread bit, position)
        for (uint32_t tb = 0; tb < M_; ++tb)
            if (thr_of_tile_[tb] >= 0) pos.push_back({thr_of_tile_[tb], global ? qpos[tb] : tb});
        std::sort(pos.begin(), pos.end());
        std::string s;
        size_t i = 0;
        while (i < pos.size()) {
            size_t j = i + 1;
            while (j < pos.size() && pos[j].first == pos[j - 1].first + 1 && pos[j].second == pos[j - 1].second + 1)
                ++j;
            const uint32_t len = uint32_t(j - i);
            std::string term = "((t >> " + std::to_string(pos[i].first) + ") & " + mask64(len) + "u)";
            if (global) term = "((unsigned long long)" + term + " << " + std::to_string(pos[i].second) + ")";
            else term = "(" + term + " << " + std::to_string(pos[i].second) + ")";
            s += s.empty() ? term : " | " + term;
            i = j;
        }
        return s.empty() ? (global ? "0ull" : "0u") : "(" + s + ")";
    }
    std::string thread_global() const { return thread_scatter(true); }

    uint64_t reg_global(uint32_t k) const { return reg_global(k, q_of_tile_); }
    uint64_t reg_global(uint32_t k, const std::vector<uint32_t>& qpos) const {
        uint64_t g = 0;
        for (uint32_t i = 0; i < RB_; ++i)
            if ((k >> i) & 1) g |= uint64_t(1) << qpos[rb_[i]];
        return g;
    }
    uint32_t reg_tile(uint32_t k) const {
        uint32_t j = 0;
        for (uint32_t i = 0; i < RB_; ++i)
            if ((k >> i) & 1) j |= 1u << rb_[i];
        return j;
    }
    // Exchange addressing.  The shared array holds the tile (in a half
    // exchange, the half with the round bit taken out) at element address
    // a(j) = j ^ (sum over index bits p >= nb of bit p of j times h[p]).  The
    // low nb bits of the address pick the bank slot (nb = 4 for 16-byte
    // elements, 5 for 8-byte ones: a 256-byte row of banks), and a() is a
    // bijection, as only low bits change, by a function of high bits.  Each
    // exchange picks its own h so that the lanes the LDS serves together
    // (lane bits 0..nb-1) hit distinct bank slots on both sides: the index
    // bits those lanes carry before and after the exchange must map to
    // independent vectors.  a() is linear, so a thread's base address and
    // each register's offset are computed apart and combined with one XOR.
    struct Swz {
        std::vector<uint32_t> h;
        bool add = false;  // thread and register parts of every address have disjoint bits
    };
    uint32_t nb() const { return dbl_ ? 4 : 5; }
    static uint32_t swz_apply(const Swz& z, uint32_t j) {
        uint32_t a = j;
        for (size_t p = 0; p < z.h.size(); ++p)
            if ((j >> p) & 1) a ^= z.h[p];
        return a;
    }
    static std::string swz_expr(const Swz& z, const std::string& j) {
        std::string s = j;
        for (size_t p = 0; p < z.h.size(); ++p) {
            const uint32_t v = z.h[p];
            if (!v) continue;
            if ((v & (v - 1)) == 0) {  // one bit b < p: move bit p of j down to b
                uint32_t b = 0;
                while (!((v >> b) & 1)) ++b;
                s += " ^ ((" + j + " >> " + std::to_string(p - b) + ") & " + std::to_string(v) + "u)";
            } else {
                s += " ^ ((0u - ((" + j + " >> " + std::to_string(p) + ") & 1u)) & " + std::to_string(v) + "u)";
            }
        }
        return "(" + s + ")";
    }
    // Index bits (after taking out bit c; c < 0: none) carried by lane bits
    // 0..nb-1 of a thread map.  A lane bit on the round bit c is constant
    // among the lanes active in a round and is left out.
    std::vector<int> lane_bits(const std::vector<int>& thr, int c) const {
        std::vector<int> pos;
        for (uint32_t u = 0; u < nb(); ++u)
            for (uint32_t t = 0; t < M_; ++t)
                if (thr[t] == int(u) && int(t) != c) pos.push_back(c >= 0 && int(t) > c ? int(t) - 1 : int(t));
        return pos;
    }
    // Rank over GF(2) of the bank-slot vectors of index bits `ps` (low `bits`
    // bits of each).
    static uint32_t rank_of(const Swz& z, const std::vector<int>& ps, uint32_t n, uint32_t bits) {
        std::vector<uint32_t> basis;
        for (int p : ps) {
            uint32_t v = (p < int(n) ? (1u << p) : z.h[size_t(p)]) & ((1u << bits) - 1);
            for (uint32_t b : basis)
                if ((v ^ b) < v) v ^= b;
            if (v) {
                basis.push_back(v);
                std::sort(basis.begin(), basis.end(), [](uint32_t x, uint32_t y) { return x > y; });
            }
        }
        return uint32_t(basis.size());
    }
    // h for one exchange between lane bit sets w (write side) and r (read
    // side) of a buffer of `bits` index bits.  Full rank on both sides is
    // required; also asking it of the first nb-1 lanes over nb-1 bank bits
    // (LDS that serves half as many lanes at once) is preferred; single-bit
    // vectors are tried first (two instructions each).  No solution: h = 0.
    // Buffer index bits (bit c taken out; c < 0: none) of the tile bits that
    // are thread bits (want_reg false) or register bits of a layout.
    std::vector<int> buffer_bits(const Layout& L, int c, bool want_reg) const {
        std::vector<int> out;
        for (uint32_t t = 0; t < M_; ++t)
            if ((want_reg ? L.reg[t] : L.thr[t]) >= 0 && int(t) != c) out.push_back(c >= 0 && int(t) > c ? int(t) - 1 : int(t));
        return out;
    }
    // Also preferred: on each side, the bits a thread's part of an address
    // can set and those a register's part can set are disjoint, so the
    // register offsets add to the thread base (the LDS instruction's
    // immediate offset) instead of taking an XOR per access.
    Swz choose_swizzle(const std::vector<int>& w, const std::vector<int>& r, uint32_t bits, const Layout& lw,
                       const Layout& lr, int c) const {
        const uint32_t n = nb();
        Swz z;
        z.h.assign(bits, 0);
        const std::vector<int> wt = buffer_bits(lw, c, false), wr = buffer_bits(lw, c, true);
        const std::vector<int> rt = buffer_bits(lr, c, false), rr = buffer_bits(lr, c, true);
        auto used = [&](const std::vector<int>& ps) {
            uint32_t u = 0;
            for (int p : ps) u |= (1u << p) | z.h[size_t(p)];
            return u;
        };
        auto additive = [&] { return (used(wt) & used(wr)) == 0 && (used(rt) & used(rr)) == 0; };
        std::vector<int> free;
        for (const std::vector<int>* side : {&w, &r})
            for (int p : *side)
                if (p >= int(n) && std::find(free.begin(), free.end(), p) == free.end()) free.push_back(p);
        std::vector<uint32_t> cand;
        for (uint32_t b = 0; b < n; ++b) cand.push_back(1u << b);
        for (uint32_t v = 1; v < (1u << n); ++v)
            if (v & (v - 1)) cand.push_back(v);
        cand.push_back(0);  // no fold: admitted only with rank slack
        // slack: lanes allowed to share a bank slot's row (rank deficit).
        int slack = 0;
        auto ok = [&](bool strict, bool add) {
            if (add && !additive()) return false;
            for (const std::vector<int>* side : {&w, &r}) {
                if (int(rank_of(z, *side, n, n)) + slack < int(side->size())) return false;
                if (strict) {
                    std::vector<int> sub;
                    for (int p : *side)
                        if (sub.size() + 1 < n) sub.push_back(p);
                    if (rank_of(z, sub, n, n - 1) < sub.size()) return false;
                }
            }
            return true;
        };
        std::function<bool(size_t, bool, bool)> search = [&](size_t i, bool strict, bool add) {
            if (i == free.size()) return ok(strict, add);
            for (uint32_t v : cand) {
                z.h[size_t(free[i])] = v;
                if (search(i + 1, strict, add)) return true;
            }
            z.h[size_t(free[i])] = 0;
            return false;
        };
        if (search(0, true, true) || search(0, false, true)) {
            z.add = true;
            return z;
        }
        // An add per access costs no instruction: it is worth a two-way bank
        // conflict on one side (measured; more sharing than that is not).
        slack = 1;
        if (search(0, false, true)) {
            z.add = true;
            return z;
        }
        slack = 0;
        if (!search(0, true, false) && !search(0, false, false)) z.h.assign(bits, 0);
        z.add = additive();
        return z;
    }

    // a *= (fr + i fi) (fi empty: real factor).
    void acc_mul(Acc& a, const std::string& fr, const std::string& fi) {
        if (!a.on) {
            o_ << "  " << a.r << " = " << fr << "; " << a.i << " = " << (fi.empty() ? "R(0)" : fi) << ";\n";
            a.on = true;
            a.real = fi.empty();
            return;
        }
        if (fi.empty()) {
            o_ << "  " << a.r << " = " << a.r << " * " << fr << ";";
            if (!a.real) o_ << " " << a.i << " = " << a.i << " * " << fr << ";";
            o_ << "\n";
        } else if (a.real) {
            o_ << "  " << a.i << " = " << a.r << " * " << fi << "; " << a.r << " = " << a.r << " * " << fr << ";\n";
            a.real = false;
        } else {
            o_ << "  { const R x = " << a.r << "; " << a.r << " = " << a.r << " * " << fr << " - " << a.i << " * " << fi
               << "; " << a.i << " = x * " << fi << " + " << a.i << " * " << fr << "; }\n";
        }
    }
    void acc_mul_lit(Acc& a, double fr, double fi) {
        if (fr == 1.0 && fi == 0.0) return;
        acc_mul(a, hexlit(fr), fi == 0.0 ? "" : hexlit(fi));
    }
    // Applies accumulator a to the registers k with (k >> bit & 1) == v
    // (bit < 0: all registers) and clears it.
    void acc_apply(Acc& a, int bit, int v) {
        if (!a.on) return;
        for (uint32_t k = 0; k < NR_; ++k)
            if (bit < 0 || int((k >> bit) & 1) == v) mul_rt(k, a.r, a.real ? "" : a.i);
        a.on = false;
        a.real = true;
    }
    void acc_flush_bit(uint32_t b) {
        acc_apply(acc_[1 + 2 * b], int(b), 0);
        acc_apply(acc_[2 + 2 * b], int(b), 1);
    }
    void acc_flush_all() {
        acc_apply(acc_[0], -1, 0);
        for (uint32_t b = 0; b < RB_; ++b) acc_flush_bit(b);
    }

    // Half-size exchanges: an exchange can go through a buffer of half the
    // tile in two rounds when some tile bit c is the same thread bit in the
    // phase before and the phase after: in round h the threads whose bit c is
    // h write their amplitudes (all with bit c = h) and read back theirs, each
    // thread reading only what it will hold, so no register is overwritten
    // before it is written out.  A wave-level bit (thread bit >= 6) keeps
    // whole waves together.  The kernel then needs half the LDS, and more
    // workgroups fit on a CU.  An exchange whose two layouts share no thread
    // bit goes through an intermediate layout that shares one with each.
    void plan_half_exchanges() {
        half_ = false;
        half_x_.clear();
        // Two-state (adjoint) tiles hold twice the data: half buffers from 10
        // qubits.
        if (M_ - RB_ <= 6 || M_ < (ns_ == 2 ? 10u : 11u)) return;
        const char* env = std::getenv("QF_JIT_HALF_LDS");
        if (env && env[0] == '0') return;
        std::vector<Layout> seq;
        for (const TileOp& op : p_.ops)
            if (op.kind == TOP_PHASE) seq.push_back(layout_of(op));
        if (relabel_) seq.push_back(store_layout());
        for (size_t x = 1; x < seq.size(); ++x) {
            const Layout& a = seq[x - 1];
            const Layout& b = seq[x];
            HalfX h;
            h.c = common_thread_bit(a.thr, b.thr);
            if (h.c < 0) {
                // Through an intermediate layout: b's, with the thread bits of
                // b at tile bits c and d exchanged so that c (a thread bit in
                // both a and b, the highest in a) keeps its thread bit from a.
                int c = -1;
                for (uint32_t t = 0; t < M_; ++t)
                    if (a.thr[t] >= 0 && b.thr[t] >= 0 && (c < 0 || a.thr[t] > a.thr[c])) c = int(t);
                if (c < 0) return;
                int d = -1;
                for (uint32_t t = 0; t < M_; ++t)
                    if (b.thr[t] == a.thr[c]) d = int(t);
                if (d < 0) return;
                h.mid = b;
                h.mid.thr[c] = a.thr[c];
                h.mid.thr[d] = b.thr[c];
                h.hop = true;
                h.c_mid = c;
                h.c = common_thread_bit(h.mid.thr, b.thr);
                if (h.c < 0) return;
            }
            half_x_.push_back(h);
        }
        half_ = !half_x_.empty();
    }

    // The tile bit with the highest thread bit that is the same thread bit in
    // both maps (-1: none).  A lane bit (< 6) splits every wave into two
    // rounds, which costs twice the LDS instructions of that exchange but
    // keeps the half buffer.
    int common_thread_bit(const std::vector<int>& a, const std::vector<int>& b) const {
        int c = -1, best = -1;
        for (uint32_t tb = 0; tb < M_; ++tb)
            if (a[tb] >= 0 && a[tb] == b[tb] && a[tb] > best) {
                best = a[tb];
                c = int(tb);
            }
        return c;
    }

    void exchange(const Layout& next, bool lds_used) {
        if (half_) {
            const HalfX& h = half_x_.at(xi_++);
            if (h.hop) {
                exchange_half(h.mid, lds_used, h.c_mid);
                lds_used = true;
            }
            exchange_half(next, lds_used, h.c);
            return;
        }
        acc_flush_all();
        const Swz z = choose_swizzle(lane_bits(thr_of_tile_, -1), lane_bits(next.thr, -1), M_,
                                     Layout{reg_of_tile_, thr_of_tile_}, next, -1);
        if (lds_used) o_ << "  __syncthreads();\n";
        o_ << "  {\n    const unsigned jt = " << thread_scatter(false) << ";\n";
        o_ << "    const unsigned st = " << swz_expr(z, "jt") << " << " << csz_log() << ";\n";
        // Lambda's tile follows psi's in the shared array.
        each_state([&] {
            for (uint32_t k = 0; k < NR_; ++k)
                o_ << "    " << lds_at("st", swz_apply(z, reg_tile(k)) | (uint32_t(cur_) << M_), z.add) << " = C{" << rv(k)
                   << ", " << iv(k) << "};\n";
        });
        o_ << "  }\n  __syncthreads();\n";
        set_layout(next);
        for (uint32_t k = 0; k < NR_; ++k) var_[k] = int(k);
        o_ << "  {\n    const unsigned jt = " << thread_scatter(false) << ";\n";
        o_ << "    const unsigned st = " << swz_expr(z, "jt") << " << " << csz_log() << ";\n";
        for (int sn = 0; sn < ns_; ++sn)
            for (uint32_t k = 0; k < NR_; ++k)
                o_ << "    { const C v = " << lds_at("st", swz_apply(z, reg_tile(k)) | (uint32_t(sn) << M_), z.add) << "; "
                   << reg_name(sn, 'r', k) << " = v.re; " << reg_name(sn, 'i', k) << " = v.im; }\n";
        o_ << "  }\n";
    }

    // The shared array's element at byte offset `base` ^ (element address
    // a * element size): one XOR per access on a byte offset computed once
    // per exchange, or, when the two parts have disjoint bits, an add that
    // the LDS instruction's immediate offset takes.
    uint32_t csz_log() const { return dbl_ ? 4 : 3; }
    std::string lds_at(const std::string& base, uint32_t j, bool add = false) const {
        return "(*(C*)((char*)xb + (" + base + (add ? " + " : " ^ ") + std::to_string(j << csz_log()) + "u)))";
    }

    // Tile index with bit c taken out (the half-buffer index).
    static uint32_t drop_bit(uint32_t j, uint32_t c) { return (j & ((1u << c) - 1)) | ((j >> (c + 1)) << c); }
    std::string drop_expr(const std::string& j, uint32_t c) const {
        return "((" + j + " & " + std::to_string((1u << c) - 1) + "u) | ((" + j + " >> " + std::to_string(c + 1) +
               ") << " + std::to_string(c) + "))";
    }

    void exchange_half(const Layout& next, bool lds_used, int c) {
        acc_flush_all();
        const uint32_t cb = uint32_t(c);
        const Swz z = choose_swizzle(lane_bits(thr_of_tile_, c), lane_bits(next.thr, c), M_ - 1,
                                     Layout{reg_of_tile_, thr_of_tile_}, next, c);
        std::vector<std::string> writes;
        each_state([&] {
            for (uint32_t k = 0; k < NR_; ++k)
                writes.push_back(lds_at("sw", swz_apply(z, drop_bit(reg_tile(k), cb)) | (uint32_t(cur_) << (M_ - 1)), z.add) +
                                 " = C{" + rv(k) + ", " + iv(k) + "};");
        });
        o_ << "  {\n    const unsigned jw = " << thread_scatter(false) << ";\n";
        o_ << "    const unsigned dw = " << drop_expr("jw", cb) << ";\n";
        o_ << "    const unsigned sw = " << swz_expr(z, "dw") << " << " << csz_log() << ";\n";
        set_layout(next);
        for (uint32_t k = 0; k < NR_; ++k) var_[k] = int(k);
        o_ << "    const unsigned jr = " << thread_scatter(false) << ";\n";
        o_ << "    const unsigned dr = " << drop_expr("jr", cb) << ";\n";
        o_ << "    const unsigned sr = " << swz_expr(z, "dr") << " << " << csz_log() << ";\n";
        o_ << "    const unsigned h = (jw >> " << cb << ") & 1u;\n";
        for (int round = 0; round < 2; ++round) {
            if (lds_used || round > 0) o_ << "    __syncthreads();\n";
            o_ << "    if (h == " << round << "u) {\n";
            for (const std::string& w : writes) o_ << "      " << w << "\n";
            o_ << "    }\n    __syncthreads();\n";
            o_ << "    if (h == " << round << "u) {\n";
            for (int sn = 0; sn < ns_; ++sn)
                for (uint32_t k = 0; k < NR_; ++k)
                    o_ << "      { const C v = "
                       << lds_at("sr", swz_apply(z, drop_bit(reg_tile(k), cb)) | (uint32_t(sn) << (M_ - 1)), z.add) << "; "
                       << reg_name(sn, 'r', k) << " = v.re; " << reg_name(sn, 'i', k) << " = v.im; }\n";
            o_ << "    }\n";
        }
        o_ << "  }\n";
    }

    Bit bit_of(int g) const {
        const int tb = tile_of_q_[g];
        if (tb < 0) return Bit{Bit::BLOCK, g};
        if (reg_of_tile_[tb] >= 0) return Bit{Bit::REG, reg_of_tile_[tb]};
        return Bit{Bit::THREAD, thr_of_tile_[tb]};
    }
    std::string bit_expr(const Bit& b) const {
        if (b.kind == Bit::THREAD) return "((t >> " + std::to_string(b.idx) + ") & 1u)";
        return "((unsigned)(base >> " + std::to_string(b.idx) + ") & 1u)";
    }

    // Emits f() once per carried state (rv()/iv() name that state's
    // registers).
    template <class F>
    void each_state(F f) {
        for (cur_ = 0; cur_ < ns_; ++cur_) f();
        cur_ = 0;
    }

    // a_k *= (dr + i di), literal, in every state.
    void mul_lit(uint32_t k, double dr, double di) {
        each_state([&] { mul_lit1(k, dr, di); });
    }
    void mul_lit1(uint32_t k, double dr, double di) {
        const std::string r = rv(k), i = iv(k);
        if (dr == 1.0 && di == 0.0) return;
        if (di == 0.0) {
            o_ << "  " << r << " = " << lincomb({lit_term(dr, r)}) << "; " << i << " = " << lincomb({lit_term(dr, i)})
               << ";\n";
            return;
        }
        if (dr == 0.0) {
            o_ << "  { const R x = " << r << "; " << r << " = " << lincomb({lit_term(-di, i)}) << "; " << i << " = "
               << lincomb({lit_term(di, "x")}) << "; }\n";
            return;
        }
        o_ << "  { const R x = " << r << "; " << r << " = " << lincomb({lit_term(dr, r), lit_term(-di, i)}) << "; "
           << i << " = " << lincomb({lit_term(dr, i), lit_term(di, "x")}) << "; }\n";
    }
    // a_k *= (DR + i DI), run-time values (DI empty: real factor), in every
    // state.
    void mul_rt(uint32_t k, const std::string& DR, const std::string& DI, bool neg_im = false) {
        each_state([&] { mul_rt1(k, DR, DI, neg_im); });
    }
    void mul_rt1(uint32_t k, const std::string& DR, const std::string& DI, bool neg_im) {
        const std::string r = rv(k), i = iv(k);
        if (DI.empty()) {
            o_ << "  " << r << " = " << DR << " * " << r << "; " << i << " = " << DR << " * " << i << ";\n";
            return;
        }
        o_ << "  { const R x = " << r << "; " << r << " = " << lincomb({{DR, false, r}, {DI, !neg_im, i}}) << "; " << i
           << " = " << lincomb({{DR, false, i}, {DI, neg_im, "x"}}) << "; }\n";
    }

    void swap_vars(uint32_t a, uint32_t b) { std::swap(var_[a], var_[b]); }

    // Literal ops carry a scale out of their matrix.  A dense one divided by
    // one of its entries' magnitudes sigma has entries of magnitude 1, which
    // need no multiply (Rx and Ry become one FMA per output component
    // instead of a multiply and an FMA); a diagonal one divided by its first
    // entry leaves that entry's registers untouched (Rz multiplies half the
    // registers).  The true amplitudes are then g times the tile's; g is
    // global to the tile (every dense op touches every amplitude once,
    // everything else is linear), so the next literal dense op takes it into
    // its matrix and the store applies what is left.
    void fold_scale(double* m, int count) {
        if (g_ == zc(1.0, 0.0)) return;
        for (int x = 0; x + 1 < count; x += 2) {
            const zc v = zc(m[x], m[x + 1]) * g_;
            m[x] = v.real();
            m[x + 1] = v.imag();
        }
        g_ = 1.0;
    }
    // Divides the complex values d[0..n) by d[0] and multiplies g by it.
    void factor_first(zc* d, int n) {
        if (!std::isnormal(std::abs(d[0])) || d[0] == zc(1.0, 0.0)) return;
        const zc f = d[0];
        for (int x = 1; x < n; ++x) d[x] /= f;
        d[0] = 1.0;
        g_ *= f;
    }
    double take_scale(double* m, int count) {
        fold_scale(m, count);
        double best = 0;
        int best_n = 0;
        for (int x = 0; x < count; ++x) {
            const double a = std::fabs(m[x]);
            if (a == 0 || !std::isfinite(a)) continue;
            int n = 0;
            for (int y = 0; y < count; ++y) n += std::fabs(m[y]) == a;
            if (n > best_n || (n == best_n && a > best)) {
                best = a;
                best_n = n;
            }
        }
        if (best_n == 0 || best == 1.0 || !std::isnormal(best)) return 1.0;
        for (int x = 0; x < count; ++x) m[x] /= best;
        return best;
    }

    void mat1(const TileOp& op) {
        const uint32_t i = uint32_t(op.q0);
        double m[8];
        std::memcpy(m, op.m, sizeof m);
        const bool offz = m[2] == 0 && m[3] == 0 && m[4] == 0 && m[5] == 0;
        const bool diagz = m[0] == 0 && m[1] == 0 && m[6] == 0 && m[7] == 0;
        if (offz) {
            zc d[2] = {zc(m[0], m[1]), zc(m[6], m[7])};
            factor_first(d, 2);
            for (uint32_t k = 0; k < NR_; ++k) {
                const zc v = d[(k >> i) & 1];
                mul_lit(k, v.real(), v.imag());
            }
            return;
        }
        acc_flush_bit(i);
        if (diagz) {
            zc d[2] = {zc(m[2], m[3]), zc(m[4], m[5])};
            factor_first(d, 2);
            for (uint32_t k = 0; k < NR_; ++k)
                if (!((k >> i) & 1)) swap_vars(k, k | (1u << i));
            for (uint32_t k = 0; k < NR_; ++k) {
                const zc v = d[(k >> i) & 1];
                mul_lit(k, v.real(), v.imag());
            }
            return;
        }
        g_ = take_scale(m, 8);
        each_state([&] {
            for (uint32_t k = 0; k < NR_; ++k) {
                if ((k >> i) & 1) continue;
                const uint32_t k1 = k | (1u << i);
                const std::string rA = rv(k), iA = iv(k), rB = rv(k1), iB = iv(k1);
                o_ << "  { const R a0r = " << rA << ", a0i = " << iA << ", a1r = " << rB << ", a1i = " << iB << ";\n";
                o_ << "    " << rA << " = "
                   << lincomb({lit_term(m[0], "a0r"), lit_term(-m[1], "a0i"), lit_term(m[2], "a1r"),
                               lit_term(-m[3], "a1i")})
                   << ";\n";
                o_ << "    " << iA << " = "
                   << lincomb({lit_term(m[0], "a0i"), lit_term(m[1], "a0r"), lit_term(m[2], "a1i"),
                               lit_term(m[3], "a1r")})
                   << ";\n";
                o_ << "    " << rB << " = "
                   << lincomb({lit_term(m[4], "a0r"), lit_term(-m[5], "a0i"), lit_term(m[6], "a1r"),
                               lit_term(-m[7], "a1i")})
                   << ";\n";
                o_ << "    " << iB << " = "
                   << lincomb({lit_term(m[4], "a0i"), lit_term(m[5], "a0r"), lit_term(m[6], "a1i"),
                               lit_term(m[7], "a1r")})
                   << "; }\n";
            }
        });
    }

    // Dense 4x4 on register bits q0 < q1 over the index bit(q0) + 2 bit(q1),
    // literal coefficients; zero entries (most of a channel superoperator's)
    // are dropped.
    void mat2(const TileOp& op) {
        const uint32_t i = uint32_t(op.q0), j = uint32_t(op.q1);
        acc_flush_bit(i);
        acc_flush_bit(j);
        double m[32];
        std::memcpy(m, op.m, sizeof m);
        g_ = take_scale(m, 32);
        each_state([&] { mat2_regs(i, j, m); });
    }
    void mat2_regs(uint32_t i, uint32_t j, const double* m) {
        for (uint32_t k = 0; k < NR_; ++k) {
            if (((k >> i) & 1) || ((k >> j) & 1)) continue;
            const uint32_t idx[4] = {k, k | (1u << i), k | (1u << j), k | (1u << i) | (1u << j)};
            o_ << "  { const R v0r = " << rv(idx[0]) << ", v0i = " << iv(idx[0]) << ", v1r = " << rv(idx[1])
               << ", v1i = " << iv(idx[1]) << ", v2r = " << rv(idx[2]) << ", v2i = " << iv(idx[2])
               << ", v3r = " << rv(idx[3]) << ", v3i = " << iv(idx[3]) << ";\n";
            for (int r = 0; r < 4; ++r) {
                std::vector<Term> re, im;
                for (int c = 0; c < 4; ++c) {
                    const double mr = m[2 * (4 * r + c)], mi = m[2 * (4 * r + c) + 1];
                    const std::string vr = "v" + std::to_string(c) + "r", vi = "v" + std::to_string(c) + "i";
                    re.push_back(lit_term(mr, vr));
                    re.push_back(lit_term(-mi, vi));
                    im.push_back(lit_term(mr, vi));
                    im.push_back(lit_term(mi, vr));
                }
                o_ << "    " << rv(idx[r]) << " = " << lincomb(re) << "; " << iv(idx[r]) << " = " << lincomb(im)
                   << ";\n";
            }
            o_ << "  }\n";
        }
    }

    // Diagonal d[bit(g1) * 2 + bit(g0)] (g1 < 0: one qubit, d[bit(g0)]).
    void diag(const TileOp& op) {
        const int g0 = op.q0, g1 = op.q1;
        const bool two = g1 >= 0;
        const Bit b0 = bit_of(g0);
        const Bit b1 = two ? bit_of(g1) : Bit{Bit::REG, -1};
        auto d = [&](int idx, int c) { return op.m[2 * idx + c]; };
        const std::string pfx = "d" + std::to_string(opn_);
        const bool rt0 = b0.kind != Bit::REG, rt1 = two && b1.kind != Bit::REG;
        const int nreg = (rt0 ? 0 : 1) + (two && !rt1 ? 1 : 0);
        if (nreg == 2 || (nreg == 1 && !rt0 && !rt1)) {
            // Only register bits: a literal factor per register, the first
            // value taken out into the tile's scale.
            zc dv[4];
            for (int x = 0; x < (two ? 4 : 2); ++x) dv[x] = zc(d(x, 0), d(x, 1));
            factor_first(dv, two ? 4 : 2);
            for (uint32_t k = 0; k < NR_; ++k) {
                const int x0 = int((k >> b0.idx) & 1), x1 = two ? int((k >> b1.idx) & 1) : 0;
                const zc v = dv[x1 * 2 + x0];
                mul_lit(k, v.real(), v.imag());
            }
            return;
        }
        if (rt0) o_ << "  const bool " << pfx << "c0 = " << bit_expr(b0) << ";\n";
        if (rt1) o_ << "  const bool " << pfx << "c1 = " << bit_expr(b1) << ";\n";
        // Factor for a fixed value of the (at most one) register bit, as a
        // select among literals over the run-time bits.
        auto factor = [&](Acc& a, int xreg) {
            const int x0 = rt0 ? -1 : xreg, x1 = !two ? 0 : (rt1 ? -1 : xreg);
            std::vector<int> cand;
            for (int v1 = 0; v1 < 2; ++v1)
                for (int v0 = 0; v0 < 2; ++v0) {
                    if (x0 >= 0 && v0 != x0) continue;
                    if (!two && v1) continue;
                    if (two && x1 >= 0 && v1 != x1) continue;
                    cand.push_back(v1 * 2 + v0);
                }
            bool same = true, real = true;
            for (int c : cand) {
                if (d(c, 0) != d(cand[0], 0) || d(c, 1) != d(cand[0], 1)) same = false;
                if (d(c, 1) != 0.0) real = false;
            }
            if (same) {
                acc_mul_lit(a, d(cand[0], 0), d(cand[0], 1));
                return;
            }
            // Two distinct values: one select on the set of run-time bit
            // combinations giving the second value.  More: a product of
            // single-select factors (nested selects over the same conditions
            // become divergent branches in the compiled code).
            auto idx = [&](int v0, int v1) {
                const int i0 = x0 >= 0 ? x0 : v0, i1 = !two ? 0 : (x1 >= 0 ? x1 : v1);
                return i1 * 2 + i0;
            };
            auto eq = [&](int a, int b) { return d(a, 0) == d(b, 0) && d(a, 1) == d(b, 1); };
            auto lit = [&](double v) { return hexlit(v); };
            std::vector<std::pair<int, int>> combos;  // run-time (v0, v1)
            for (int v1 = 0; v1 < (rt1 ? 2 : 1); ++v1)
                for (int v0 = 0; v0 < (rt0 ? 2 : 1); ++v0) combos.push_back({v0, v1});
            const int base = idx(0, 0);
            int other = -1;
            bool two_values = true;
            for (auto& cb : combos) {
                const int j = idx(cb.first, cb.second);
                if (eq(j, base)) continue;
                if (other < 0) other = j;
                else if (!eq(j, other)) two_values = false;
            }
            auto cond_of = [&](int v0, int v1) {
                std::string c;
                if (rt0) c += std::string(v0 ? "" : "!") + pfx + "c0";
                if (rt1) c += std::string(c.empty() ? "" : " && ") + (v1 ? "" : "!") + pfx + "c1";
                return "(" + c + ")";
            };
            const std::string nm = pfx + "_" + std::to_string(xreg + 1);
            if (two_values) {
                std::string cond;
                for (auto& cb : combos)
                    if (!eq(idx(cb.first, cb.second), base))
                        cond += (cond.empty() ? "" : " || ") + cond_of(cb.first, cb.second);
                o_ << "  const bool " << nm << "s = " << cond << ";\n";
                o_ << "  const R " << nm << "r = " << nm << "s ? " << lit(d(other, 0)) << " : " << lit(d(base, 0)) << ";\n";
                if (!real)
                    o_ << "  const R " << nm << "i = " << nm << "s ? " << lit(d(other, 1)) << " : " << lit(d(base, 1))
                       << ";\n";
                acc_mul(a, nm + "r", real ? std::string() : nm + "i");
                return;
            }
            // Both bits run-time, three or four values: d00 * (x0 ? d01/d00 : 1)
            // * (x1 ? d10/d00 : 1) * (x0 x1 ? d11 d00 / (d01 d10) : 1).
            typedef std::complex<double> Z;
            const Z z00(d(0, 0), d(0, 1)), z01(d(1, 0), d(1, 1)), z10(d(2, 0), d(2, 1)), z11(d(3, 0), d(3, 1));
            const Z f1 = z01 / z00, f2 = z10 / z00, f3 = z11 * z00 / (z01 * z10);
            acc_mul_lit(a, z00.real(), z00.imag());
            const Z fs[3] = {f1, f2, f3};
            const char* conds[3] = {"c0", "c1", ""};
            for (int q = 0; q < 3; ++q) {
                if (fs[q] == Z(1, 0)) continue;
                const std::string fn = nm + "f" + std::to_string(q);
                const std::string cnd = q < 2 ? pfx + conds[q] : "(" + pfx + "c0 && " + pfx + "c1)";
                o_ << "  const R " << fn << "r = " << cnd << " ? " << lit(fs[q].real()) << " : R(1);\n";
                o_ << "  const R " << fn << "i = " << cnd << " ? " << lit(fs[q].imag()) << " : R(0);\n";
                acc_mul(a, fn + "r", fn + "i");
            }
        };
        if (nreg == 0 && b0.kind == Bit::BLOCK && (!two || b1.kind == Bit::BLOCK)) {
            // Tile-index bits only, unit-modulus values: a phase of the whole
            // tile, added up in fixed point (2^-64 turns) from the values'
            // own phases and applied once before the store.
            const int nv = two ? 4 : 2;
            bool unit = true;
            unsigned long long K[4] = {0, 0, 0, 0};
            for (int x = 0; x < nv; ++x) {
                const double re = d(x, 0), im = d(x, 1);
                if (std::fabs(re * re + im * im - 1.0) > 1e-14) unit = false;
                const long double t = (long double)std::atan2(im, re) / (2.0L * 3.14159265358979323846264338327950288L);
                K[x] = (unsigned long long)(long long)std::llroundl(t * 0x1p63L) << 1;
            }
            if (unit) {
                auto k = [&](int x) { return "0x" + hex64(K[x]) + "ull"; };
                o_ << "  uph += " << pfx << "c0 ? ";
                if (two)
                    o_ << "(" << pfx << "c1 ? " << k(3) << " : " << k(1) << ") : (" << pfx << "c1 ? " << k(2) << " : "
                       << k(0) << ");\n";
                else
                    o_ << k(1) << " : " << k(0) << ";\n";
                uph_ = true;
                return;
            }
        }
        if (nreg == 0) {
            factor(acc_[0], -1);
            return;
        }
        const uint32_t rb = uint32_t(rt0 ? b1.idx : b0.idx);
        factor(acc_[1 + 2 * rb], 0);
        factor(acc_[2 + 2 * rb], 1);
    }

    void cnot(const TileOp& op) {
        const uint32_t ti = uint32_t(op.q1);
        const Bit c = bit_of(op.q0);
        acc_flush_bit(ti);
        if (c.kind == Bit::REG) {
            for (uint32_t k = 0; k < NR_; ++k)
                if (((k >> c.idx) & 1) && !((k >> ti) & 1)) swap_vars(k, k | (1u << ti));
            return;
        }
        const std::string cn = "x" + std::to_string(opn_);
        o_ << "  const bool " << cn << " = " << bit_expr(c) << ";\n";
        each_state([&] {
            for (uint32_t k = 0; k < NR_; ++k) {
                if ((k >> ti) & 1) continue;
                const uint32_t k1 = k | (1u << ti);
                const std::string rA = rv(k), iA = iv(k), rB = rv(k1), iB = iv(k1);
                o_ << "  { const R ur = " << rA << ", ui = " << iA << "; " << rA << " = " << cn << " ? " << rB
                   << " : " << rA << "; " << iA << " = " << cn << " ? " << iB << " : " << iA << "; " << rB << " = "
                   << cn << " ? ur : " << rB << "; " << iB << " = " << cn << " ? ui : " << iB << "; }\n";
            }
        });
    }

    void swap(const TileOp& op) {
        const uint32_t a = uint32_t(op.q0), b = uint32_t(op.q1);
        acc_flush_bit(a);
        acc_flush_bit(b);
        for (uint32_t k = 0; k < NR_; ++k)
            if (((k >> a) & 1) && !((k >> b) & 1)) swap_vars(k, k ^ (1u << a) ^ (1u << b));
    }

    // (c, s) of a rotation occurrence from the instance's angle-table row.
    void load_cs(const TileOp& op, std::string& c, std::string& s) {
        c = "c" + std::to_string(opn_);
        s = "s" + std::to_string(opn_);
        o_ << "  const R " << c << " = R(csr[" << 2 * op.slot << "]);\n";
        o_ << "  const R " << s << " = " << ((op.flags & TF_NEGATE) ? "-" : "") << "R(csr[" << 2 * op.slot + 1
           << "]);\n";
    }

    // Im <lambda| G |psi> for a recorded rotation's generator G, before the
    // rotation: each thread's share over its registers in FP64, summed over
    // the wave (DPP row operations) into the shared table qred (row r, one
    // value per wave); the tile's sum goes out after the store.  Pending
    // per-bit diagonal factors are phases common to both states and cancel
    // in conj(lambda) psi, except across the pair of an X or Y generator's
    // bit, whose factors are applied first.  The factors common to all of a
    // thread's amplitudes (the literal g of take_scale, the thread
    // accumulator with its rotation scales) are still owed by both states,
    // so the share is multiplied by their squared magnitudes.
    void record(const TileOp& op, uint32_t r, uint32_t W) {
        const std::string v = "q" + std::to_string(opn_);
        auto val = [&](int st, char part, uint32_t k) { return "double(" + reg_name(st, part, uint32_t(var_[k])) + ")"; };
        o_ << "  double " << v << " = 0.0;\n";
        if (op.kind == TOP_ROT && op.q1 != 2) {
            const uint32_t B = uint32_t(op.q0);
            acc_flush_bit(B);
            for (uint32_t k = 0; k < NR_; ++k) {
                const uint32_t kp = k ^ (1u << B);
                const std::string lr = val(1, 'r', k), li = val(1, 'i', k), pr = val(0, 'r', kp), pi = val(0, 'i', kp);
                if (op.q1 == 0)  // X: Im(conj(l) p)
                    o_ << "  " << v << " += " << lr << " * " << pi << " - " << li << " * " << pr << ";\n";
                else  // Y: -+Re(conj(l) p) by the bit of l
                    o_ << "  " << v << ((k >> B) & 1 ? " += " : " -= ") << "(" << lr << " * " << pr << " + " << li
                       << " * " << pi << ");\n";
            }
        } else {
            // Z (Z Z): Im(conj(l) p) with the sign of the generator's parity.
            std::vector<Bit> bits{bit_of(op.q0)};
            if (op.kind == TOP_ROTZZ) bits.push_back(bit_of(op.q1));
            std::string rt;
            for (const Bit& b : bits)
                if (b.kind != Bit::REG) rt += (rt.empty() ? "" : " ^ ") + bit_expr(b);
            for (uint32_t k = 0; k < NR_; ++k) {
                int par = 0;
                for (const Bit& b : bits)
                    if (b.kind == Bit::REG) par ^= int((k >> b.idx) & 1);
                o_ << "  " << v << (par ? " -= " : " += ") << "(" << val(1, 'r', k) << " * " << val(0, 'i', k) << " - "
                   << val(1, 'i', k) << " * " << val(0, 'r', k) << ");\n";
            }
            if (!rt.empty()) o_ << "  if (" << rt << ") " << v << " = -" << v << ";\n";
        }
        if (g_ != zc(1.0, 0.0)) o_ << "  " << v << " *= " << hexlit(std::norm(g_)).substr(1) << ";\n";
        if (acc_[0].on)
            o_ << "  " << v << " *= double(" << acc_[0].r << ") * double(" << acc_[0].r << ")"
               << (acc_[0].real ? "" : " + double(" + acc_[0].i + ") * double(" + acc_[0].i + ")") << ";\n";
        // Pin the share here: otherwise the compiler sinks its arithmetic to
        // the batch's wave sum and keeps the pre-rotation amplitudes alive
        // until then (twice the registers).
        o_ << "  QF_PIN(" << v << ");\n";
        if (pend_.empty()) rec0_ = r;
        recw_ = W;
        pend_.push_back(v);
        if (pend_.size() == std::min<size_t>(kRecBatch, std::min<uint32_t>(NT_, 64))) flush_records();
    }

    // The wave sums of the pending records, all at once.  Transpose steps on
    // lane bits 0 and 1 (each lane keeps one half of the list and adds its
    // partner's copy of that half) leave lane l with record l & (B - 1) of
    // the batch summed over its quad; adding the partners
    // across the higher lane bits (row rotations by 4 and 8 lanes, swizzles
    // across 16 and a permute across 32) completes every record's sum in
    // every lane of its class, and lanes 0..B-1 write them to the shared
    // table (row = record, column = wave).  Lane bits the workgroup does not
    // have are skipped; the order of the additions is fixed.
    void flush_records() {
        if (pend_.empty()) return;
        const size_t n = pend_.size();
        size_t B = 1;
        while (B < n) B <<= 1;
        const uint32_t lanes = std::min<uint32_t>(NT_, 64);
        const std::string p = "rb" + std::to_string(rec0_);
        std::vector<std::string> x;
        o_ << "  {\n    const unsigned ln = t & 63u;\n";
        for (size_t k = 0; k < B; ++k) {
            x.push_back(p + "_" + std::to_string(k));
            o_ << "    double " << x.back() << " = " << (k < n ? pend_[k] : std::string("0.0")) << ";\n";
        }
        uint32_t j = 0;
        for (; (size_t(1) << j) < B; ++j) {
            const char* ctrl = j == 0 ? "0xB1" : "0x4E";  // lane ^ 1, lane ^ 2
            const std::string h = p + "h" + std::to_string(j);
            o_ << "    const bool " << h << " = (ln >> " << j << ") & 1u;\n";
            std::vector<std::string> y;
            for (size_t k = 0; k + 1 < x.size(); k += 2) {
                y.push_back(p + "_s" + std::to_string(j) + "_" + std::to_string(k / 2));
                o_ << "    const double " << y.back() << " = (" << h << " ? " << x[k + 1] << " : " << x[k]
                   << ") + qf_dpp<" << ctrl << ", 0xf>(" << h << " ? " << x[k] << " : " << x[k + 1] << ");\n";
            }
            x = y;
        }
        std::string v = x[0];
        for (; (1u << j) < lanes; ++j) {
            const uint32_t m = 1u << j;
            const std::string w = p + "_x" + std::to_string(m);
            std::string partner;
            if (m == 1) partner = "qf_dpp<0xB1, 0xf>(" + v + ")";
            else if (m == 2) partner = "qf_dpp<0x4E, 0xf>(" + v + ")";
            else if (m < 16 && lanes >= 16) partner = "qf_dpp<" + std::to_string(0x120 + m) + ", 0xf>(" + v + ")";
            else if (m < 32) partner = "qf_swz<" + std::to_string((m << 10) | 0x1F) + ">(" + v + ")";
            else partner = "qf_bperm32(" + v + ", ln)";
            // (Row rotations by 4 then 8 add the four lanes of a class in a row.)
            o_ << "    const double " << w << " = " << v << " + " << partner << ";\n";
            v = w;
        }
        // Lane l (l < B) holds record l of the batch.
        const std::string rec = B == 1 ? std::string("0u") : "(ln & " + std::to_string(B - 1) + "u)";
        o_ << "    if (ln < " << B << "u && " << rec << " < " << n << "u) qred[(" << rec0_ << "u + " << rec << ") * "
           << recw_ << "u + (t >> 6)] = " << v << ";\n  }\n";
        pend_.clear();
    }

    void rot(const TileOp& op) {
        if (op.q1 != 2 && scaled_) {
            rot_xy_scaled(op);
            return;
        }
        std::string c, s;
        load_cs(op, c, s);
        if (op.q1 == 2) {  // exp(-i theta Z / 2): (c - is) on bit 0, (c + is) on bit 1
            const Bit b = bit_of(op.q0);
            if (b.kind == Bit::REG) {
                // (c - is) for the whole thread (its scale accumulator), and
                // (c + is) / (c - is) = (c^2 - s^2) + i 2cs on bit 1 only.
                const std::string er = "e" + std::to_string(opn_) + "r", ei = "e" + std::to_string(opn_) + "i";
                o_ << "  const R " << er << " = " << c << " * " << c << " - " << s << " * " << s << ", " << ei
                   << " = R(2) * " << c << " * " << s << ";\n";
                acc_mul(acc_[0], c, "(-" + s + ")");
                for (uint32_t k = 0; k < NR_; ++k)
                    if ((k >> b.idx) & 1) mul_rt(k, er, ei);
                return;
            }
            const std::string sz = "z" + std::to_string(opn_);
            o_ << "  const R " << sz << " = " << bit_expr(b) << " ? " << s << " : -" << s << ";\n";
            acc_mul(acc_[0], c, sz);
            return;
        }
        const uint32_t i = uint32_t(op.q0);
        acc_flush_bit(i);
        // Direct form: [[c, -s], [s, c]] (Ry) and [[c, -is], [-is, c]] (Rx)
        // with the half-angle values from the table, a multiply and an FMA
        // per output component, no division and no branch.
        each_state([&] { rot_xy_direct(op, i, c, s); });
    }
    // The row of a scaled X/Y rotation holds f, the larger of c, s in
    // magnitude, and u, the other over f, whose lowest mantissa bit is 0 for
    // f = c and 1 for f = s (the host's choice per instance).  The flags of
    // the pass's rotations are read once at the start into scalar words, so
    // each rotation branches on a bit of a word (uniform over the workgroup,
    // no wait on a load); each output component is one FMA (the forms of
    // rot_xy_regs), the ratio of the next rotation is loaded before this
    // one's branch, and f goes into the thread's scale accumulator after it.
    // For the inverse rotation (s -> -s) both forms take -u, and the f = s
    // form also -f.
    void scaled_flags() {
        for (size_t i = 1; i < p_.ops.size(); ++i) {
            const TileOp& op = p_.ops[i];
            if (op.kind == TOP_ROT && op.q1 != 2) {
                sc_pos_[i] = uint32_t(sc_ops_.size());
                sc_ops_.push_back(i);
            }
        }
        for (size_t w = 0; w * 32 < sc_ops_.size(); ++w) {
            o_ << "  const unsigned fw" << w << " = __builtin_amdgcn_readfirstlane(int(0u";
            for (size_t j = w * 32; j < std::min(sc_ops_.size(), w * 32 + 32); ++j)
                o_ << "\n      | ((unsigned(__builtin_bit_cast(unsigned long long, csr["
                   << 2 * p_.ops[sc_ops_[j]].slot + 1 << "])) & 1u) << " << j % 32 << ")";
            o_ << "));\n";
        }
    }
    const std::string& scaled_ratio(size_t i) {
        auto it = sc_u_.find(i);
        if (it != sc_u_.end()) return it->second;
        const TileOp& op = p_.ops[i];
        const std::string u = "u" + std::to_string(sc_pos_[i]) + "_" + std::to_string(opn_);
        o_ << "  const R " << u << " = " << ((op.flags & TF_NEGATE) ? "-" : "") << "R(csr[" << 2 * op.slot + 1
           << "]);\n";
        return sc_u_[i] = u;
    }
    void rot_xy_scaled(const TileOp& op) {
        const uint32_t i = uint32_t(op.q0);
        acc_flush_bit(i);
        const uint32_t pos = sc_pos_.at(cur_op_);
        const std::string u = scaled_ratio(cur_op_);
        if (pos + 1 < sc_ops_.size()) scaled_ratio(sc_ops_[pos + 1]);
        const std::string w = "((fw" + std::to_string(pos / 32) + " >> " + std::to_string(pos % 32) + ") & 1u)";
        each_state([&] {
            o_ << "  if (" << w << ") {\n";
            rot_xy_regs(op, i, u, false);
            o_ << "  } else {\n";
            rot_xy_regs(op, i, u, true);
            o_ << "  }\n";
        });
        const std::string f = "f" + std::to_string(opn_), x = "R(csr[" + std::to_string(2 * op.slot) + "])";
        o_ << "  const R " << f << " = " << ((op.flags & TF_NEGATE) ? "(" + w + " ? -" + x + " : " + x + ")" : x)
           << ";\n";
        acc_mul(acc_[0], f, "");
    }
    void rot_xy_direct(const TileOp& op, uint32_t i, const std::string& c, const std::string& s) {
        for (uint32_t k = 0; k < NR_; ++k) {
            if ((k >> i) & 1) continue;
            const uint32_t k1 = k | (1u << i);
            const std::string rA = rv(k), iA = iv(k), rB = rv(k1), iB = iv(k1);
            o_ << "  { const R a0r = " << rA << ", a0i = " << iA << ", a1r = " << rB << ", a1i = " << iB << ";\n";
            if (op.q1 == 0) {
                o_ << "    " << rA << " = " << c << " * a0r + " << s << " * a1i; " << iA << " = " << c << " * a0i - " << s
                   << " * a1r;\n";
                o_ << "    " << rB << " = " << c << " * a1r + " << s << " * a0i; " << iB << " = " << c << " * a1i - " << s
                   << " * a0r; }\n";
            } else {
                o_ << "    " << rA << " = " << c << " * a0r - " << s << " * a1r; " << iA << " = " << c << " * a0i - " << s
                   << " * a1i;\n";
                o_ << "    " << rB << " = " << s << " * a0r + " << c << " * a1r; " << iB << " = " << s << " * a0i + " << c
                   << " * a1i; }\n";
            }
        }
    }
    // Rx = c [[1, -it], [-it, 1]] = s [[u, -i], [-i, u]] and
    // Ry = c [[1, -t], [t, 1]] = s [[u, -1], [1, u]] (t = s / c, u = c / s),
    // the scale outside: by_c selects the first form, with ratio t = u.
    void rot_xy_regs(const TileOp& op, uint32_t i, const std::string& u, bool by_c) {
        for (uint32_t k = 0; k < NR_; ++k) {
            if ((k >> i) & 1) continue;
            const uint32_t k1 = k | (1u << i);
            const std::string rA = rv(k), iA = iv(k), rB = rv(k1), iB = iv(k1);
            o_ << "  { const R a0r = " << rA << ", a0i = " << iA << ", a1r = " << rB << ", a1i = " << iB << ";\n";
            if (op.q1 == 0 && by_c) {
                o_ << "    " << rA << " = a0r + " << u << " * a1i; " << iA << " = a0i - " << u << " * a1r;\n";
                o_ << "    " << rB << " = a1r + " << u << " * a0i; " << iB << " = a1i - " << u << " * a0r; }\n";
            } else if (op.q1 == 0) {
                o_ << "    " << rA << " = " << u << " * a0r + a1i; " << iA << " = " << u << " * a0i - a1r;\n";
                o_ << "    " << rB << " = " << u << " * a1r + a0i; " << iB << " = " << u << " * a1i - a0r; }\n";
            } else if (by_c) {
                o_ << "    " << rA << " = a0r - " << u << " * a1r; " << iA << " = a0i - " << u << " * a1i;\n";
                o_ << "    " << rB << " = " << u << " * a0r + a1r; " << iB << " = " << u << " * a0i + a1i; }\n";
            } else {
                o_ << "    " << rA << " = " << u << " * a0r - a1r; " << iA << " = " << u << " * a0i - a1i;\n";
                o_ << "    " << rB << " = a0r + " << u << " * a1r; " << iB << " = a0i + " << u << " * a1i; }\n";
            }
        }
    }

    // exp(-i theta Z_g0 Z_g1 / 2): (c - is) on even parity, (c + is) on odd.
    void rotzz(const TileOp& op) {
        std::string c, s;
        load_cs(op, c, s);
        const Bit b0 = bit_of(op.q0), b1 = bit_of(op.q1);
        std::string rt;
        for (const Bit& b : {b0, b1})
            if (b.kind != Bit::REG) rt += (rt.empty() ? "" : " ^ ") + bit_expr(b);
        std::string sz = s;
        if (!rt.empty()) {
            sz = "z" + std::to_string(opn_);
            o_ << "  const R " << sz << " = (" << rt << ") ? " << s << " : -" << s << ";\n";
        }
        if (b0.kind != Bit::REG && b1.kind != Bit::REG) {
            acc_mul(acc_[0], c, sz);
            return;
        }
        if ((b0.kind == Bit::REG) != (b1.kind == Bit::REG)) {
            // One register bit r: parity = run-time parity ^ bit r, so the
            // factor is (c, sz) for r = 0 and (c, -sz) for r = 1.
            const uint32_t r = uint32_t(b0.kind == Bit::REG ? b0.idx : b1.idx);
            acc_mul(acc_[1 + 2 * r], c, sz);
            acc_mul(acc_[2 + 2 * r], c, "(-" + sz + ")");
            return;
        }
        for (uint32_t k = 0; k < NR_; ++k) {
            int par = 0;
            if (b0.kind == Bit::REG) par ^= int((k >> b0.idx) & 1);
            if (b1.kind == Bit::REG) par ^= int((k >> b1.idx) & 1);
            // Without run-time bits the sign is fixed: odd -> +s, even -> -s.
            // With them, sz already carries the run-time parity; a set
            // compile-time parity flips it.
            const bool neg = rt.empty() ? !par : par;
            mul_rt(k, c, sz, neg);
        }
    }
};

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

std::string cache_dir() {
    const char* e = std::getenv("QF_JIT_CACHE");
    if (e && (!std::strcmp(e, "0") || !std::strcmp(e, "off") || !*e)) return "";
    if (e) return e;
    const char* home = std::getenv("HOME");
    if (!home || !*home) return "";
    return std::string(home) + "/.cache/quantumforge_amd/jit";
}

void make_dirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i < path.size(); ++i) {
        cur += path[i];
        if (path[i] == '/' && cur.size() > 1) ::mkdir(cur.c_str(), 0755);
    }
    ::mkdir(path.c_str(), 0755);
}

bool read_file(const std::string& path, std::vector<char>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !out.empty();
}

void write_file_atomic(const std::string& dir, const std::string& name, const std::vector<char>& data) {
    if (dir.empty()) return;
    make_dirs(dir);
    const std::string tmp = dir + "/." + name + "." + std::to_string(::getpid()) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(data.data(), std::streamsize(data.size()));
        if (!f) return;
    }
    std::rename(tmp.c_str(), (dir + "/" + name).c_str());
}

// At most this many code objects stay in the disk cache; older ones (by
// modification time) are removed when a new program is written.
constexpr size_t kDiskCacheFiles = 4096;

void prune_disk_cache(const std::string& dir) {
    DIR* d = ::opendir(dir.c_str());
    if (!d) return;
    std::vector<std::pair<long long, std::string>> files;
    while (dirent* e = ::readdir(d)) {
        const std::string name = e->d_name;
        if (name.size() < 4 || name.compare(name.size() - 3, 3, ".co") != 0 || name[0] == '.') continue;
        struct stat st {};
        if (::stat((dir + "/" + name).c_str(), &st) == 0) files.emplace_back((long long)st.st_mtime, name);
    }
    ::closedir(d);
    if (files.size() <= kDiskCacheFiles) return;
    std::sort(files.begin(), files.end());
    for (size_t i = 0; i + kDiskCacheFiles < files.size(); ++i) ::unlink((dir + "/" + files[i].second).c_str());
}

// The compiler's version, part of every disk-cache file name: code objects
// of another hiprtc are never loaded.
std::string rtc_tag() {
    int major = 0, minor = 0;
    if (hiprtcVersion(&major, &minor) != HIPRTC_SUCCESS) return "rtc-unknown";
    return "rtc" + std::to_string(major) + "." + std::to_string(minor);
}

std::string compile_one(const std::string& src, const std::string& fname, std::vector<char>& code) {
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), fname.c_str(), 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) return std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
    const std::vector<std::string> opts = jit_compile_options();
    std::vector<const char*> argv;
    for (const auto& s : opts) argv.push_back(s.c_str());
    r = hiprtcCompileProgram(prog, int(argv.size()), argv.data());
    std::string log;
    size_t logn = 0;
    if (hiprtcGetProgramLogSize(prog, &logn) == HIPRTC_SUCCESS && logn > 1) {
        log.resize(logn);
        hiprtcGetProgramLog(prog, &log[0]);
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return std::string("hiprtcCompileProgram: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 4000);
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return "";
}

std::mutex g_mu;
std::map<std::string, std::shared_ptr<JitProgram>> g_cache;
// Insertion order of g_cache: beyond kProgramCacheEntries programs the oldest
// leaves both maps (a plan that still holds it keeps it alive).
std::deque<std::string> g_cache_order;
constexpr size_t kProgramCacheEntries = 256;
// Programs by plan fingerprint: a plan rebuilt for every call (the gradient
// engine's) finds its program without generating and hashing the source.
std::map<std::string, std::shared_ptr<JitProgram>> g_fp_cache;

// 128 bits over everything the generated source depends on: qubit count,
// precision, state count, compile options, the half-buffer switch, and every
// pass's shape, tile bits, store layout and ops.
std::string plan_fingerprint(const Plan& plan, int precision, int ns) {
    uint64_t h1 = 1469598103934665603ull, h2 = 0x9e3779b97f4a7c15ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h1 = (h1 ^ b[i]) * 1099511628211ull;
            h2 = (h2 + b[i] + 1) * 0xff51afd7ed558ccdull;
            h2 ^= h2 >> 31;
        }
    };
    auto val = [&](int64_t v) { mix(&v, sizeof v); };
    val(plan.qubits);
    val(precision);
    val(ns);
    val(plan.scaled_rotations);
    const char* half = std::getenv("QF_JIT_HALF_LDS");
    val(half && half[0] == '0' ? 0 : 1);
    for (const auto& o : jit_compile_options()) {
        mix(o.data(), o.size());
        val(-1);
    }
    for (const Pass& p : plan.passes) {
        val(p.kind);
        val(p.tile_bits);
        val(p.low);
        val(p.reg_bits);
        val(int64_t(p.high.size()));
        for (uint32_t h : p.high) val(h);
        val(int64_t(p.store_phys.size()));
        for (uint32_t h : p.store_phys) val(h);
        val(int64_t(p.ops.size()));
        for (const TileOp& o : p.ops) {
            val(o.kind);
            val(o.q0);
            val(o.q1);
            val(o.flags);
            val(o.param);
            val(o.slot);
            mix(o.m, sizeof o.m);
        }
    }
    char b[40];
    std::snprintf(b, sizeof b, "%016llx%016llx", (unsigned long long)h1, (unsigned long long)h2);
    return b;
}

}  // namespace

std::vector<std::string> jit_compile_options() {
    return {"--offload-arch=gfx950", "-O3", "-ffp-contract=fast", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops"};
}

std::string jit_pass_source(const Pass& p, uint32_t qubits, int precision, const std::string& name,
                            std::string* err, int ns, bool scaled) {
    PassGen g(p, qubits, precision, name, ns, scaled);
    if (!g.run(err)) return "";
    return g.src;
}

std::string jit_plan_source(const Plan& plan, int precision, JitProgram* prog, std::string* err, int ns) {
    // Generate every compiled pass with placeholder names; the source text
    // (with the compile options) is the cache key, and the final kernel names
    // carry the key.
    std::vector<std::string> bodies;
    prog->kernel_of_pass.assign(plan.passes.size(), -1);
    prog->threads.clear();
    prog->states = ns;
    for (size_t i = 0; i < plan.passes.size(); ++i) {
        const Pass& p = plan.passes[i];
        if (p.kind != Pass::TILE || p.reg_bits == 0) continue;
        std::string e;
        std::string body =
            jit_pass_source(p, plan.qubits, precision, "qf_pass_@" + std::to_string(bodies.size()), &e, ns,
                            plan.scaled_rotations != 0);
        if (body.empty()) {
            if (err) *err = "pass " + std::to_string(i) + ": " + e;
            return "";
        }
        prog->kernel_of_pass[i] = int(bodies.size());
        prog->threads.push_back(1u << (p.tile_bits - p.reg_bits));
        bodies.push_back(std::move(body));
    }
    // qf_wave_sum: a wave's sum in lane 63 by DPP row operations on the two
    // halves of each double (quad swaps, half-row and row mirrors, then the
    // row 15 and 31 broadcasts into rows 1, 3 and 2, 3), with no LDS traffic.
    // Defined only where the adjoint form uses it.
    std::string prelude = precision == 0 ? "struct alignas(16) C { double re, im; };\n"
                                         : "struct alignas(8) C { float re, im; };\n";
    if (ns == 2)
        prelude +=
            "#ifndef QF_WAVE_SUM_DEFINED\n"
            "template <int CTRL, int ROWS> __device__ inline double qf_dpp(double v) {\n"
            "  const long long b = __builtin_bit_cast(long long, v);\n"
            "  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xf, false);\n"
            "  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xf, false);\n"
            "  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));\n"
            "}\n"
            "#define QF_PIN(v) asm volatile(\"\" : \"+v\"(v))\n"
            "template <int PAT> __device__ inline double qf_swz(double v) {\n"
            "  const long long b = __builtin_bit_cast(long long, v);\n"
            "  const int lo = __builtin_amdgcn_ds_swizzle(int(b), PAT);\n"
            "  const int hi = __builtin_amdgcn_ds_swizzle(int(b >> 32), PAT);\n"
            "  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));\n"
            "}\n"
            "__device__ inline double qf_bperm32(double v, unsigned ln) {\n"
            "  const long long b = __builtin_bit_cast(long long, v);\n"
            "  const int a = int((ln ^ 32u) << 2);\n"
            "  const int lo = __builtin_amdgcn_ds_bpermute(a, int(b));\n"
            "  const int hi = __builtin_amdgcn_ds_bpermute(a, int(b >> 32));\n"
            "  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));\n"
            "}\n"
            "__device__ inline double qf_wave_sum(double v) {\n"
            "  v += qf_dpp<0xB1, 0xf>(v);\n  v += qf_dpp<0x4E, 0xf>(v);\n  v += qf_dpp<0x141, 0xf>(v);\n"
            "  v += qf_dpp<0x140, 0xf>(v);\n  v += qf_dpp<0x142, 0xa>(v);\n  v += qf_dpp<0x143, 0xc>(v);\n"
            "  return v;\n}\n#endif\n";
    std::string all = prelude;
    for (const auto& b : bodies) all += b;
    std::string optstr;
    for (const auto& o : jit_compile_options()) optstr += o + " ";
    char keyb[32];
    std::snprintf(keyb, sizeof keyb, "%016llx", (unsigned long long)fnv1a(optstr, fnv1a(all)));
    prog->key = keyb;
    prog->names.clear();
    prog->bodies.clear();
    for (size_t i = 0; i < bodies.size(); ++i) {
        prog->names.push_back("qf_pass_" + prog->key + "_" + std::to_string(i));
        const std::string tag = "qf_pass_@" + std::to_string(i);
        bodies[i].replace(bodies[i].find(tag), tag.size(), prog->names[i]);
        prog->bodies.push_back(std::move(bodies[i]));
    }
    prog->prelude = prelude;
    std::string out = prelude;
    for (const auto& b : prog->bodies) out += b;
    return out;
}

std::shared_ptr<JitProgram> jit_compile_plan(const Plan& plan, int precision, std::string* err, int ns) {
    const std::string fp = plan_fingerprint(plan, precision, ns);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_fp_cache.find(fp);
        if (it != g_fp_cache.end()) return it->second;
    }
    auto prog = std::make_shared<JitProgram>();
    std::string gen_err;
    if (jit_plan_source(plan, precision, prog.get(), &gen_err, ns).empty()) {
        if (err) *err = gen_err;
        return nullptr;
    }
    const std::string key = prog->key;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            g_fp_cache.emplace(fp, it->second);
            return it->second;
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    // Chunks of 8 kernels, compiled by a few host threads; each chunk is its
    // own code object (kernel names are unique).
    const size_t nk = prog->bodies.size();
    const size_t nchunk = (nk + 7) / 8;
    prog->codes.assign(nchunk, {});
    prog->code_of_kernel.assign(nk, 0);
    std::vector<std::string> srcs(nchunk, prog->prelude);
    for (size_t i = 0; i < nk; ++i) {
        const size_t c = i / 8;
        prog->code_of_kernel[i] = int(c);
        srcs[c] += prog->bodies[i];
    }
    const std::string dir = cache_dir();
    const std::string file_key = key + "_" + rtc_tag();
    bool cached = !dir.empty() && nchunk > 0;
    for (size_t c = 0; c < nchunk && cached; ++c)
        cached = read_file(dir + "/" + file_key + "_" + std::to_string(c) + ".co", prog->codes[c]);
    if (cached) {
        prog->from_disk = true;
    } else if (nchunk > 0) {
        std::vector<std::string> errs(nchunk);
        std::atomic<size_t> next{0};
        const size_t hw = std::max<size_t>(1, std::min<size_t>(8, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (size_t w = 0; w < std::min(hw, nchunk); ++w)
            pool.emplace_back([&] {
                for (size_t c = next++; c < nchunk; c = next++)
                    errs[c] = compile_one(srcs[c], "qf_pass_" + key + "_" + std::to_string(c) + ".hip", prog->codes[c]);
            });
        for (auto& th : pool) th.join();
        for (size_t c = 0; c < nchunk; ++c)
            if (!errs[c].empty()) {
                if (err) *err = errs[c];
                return nullptr;
            }
        for (size_t c = 0; c < nchunk; ++c)
            write_file_atomic(dir, file_key + "_" + std::to_string(c) + ".co", prog->codes[c]);
        if (!dir.empty()) prune_disk_cache(dir);
    }
    prog->compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::lock_guard<std::mutex> lk(g_mu);
    auto ins = g_cache.emplace(key, prog);
    if (ins.second) {
        g_cache_order.push_back(key);
        while (g_cache_order.size() > kProgramCacheEntries) {
            auto old = g_cache.find(g_cache_order.front());
            if (old != g_cache.end()) {
                for (auto f = g_fp_cache.begin(); f != g_fp_cache.end();)
                    f = f->second == old->second ? g_fp_cache.erase(f) : std::next(f);
                g_cache.erase(old);
            }
            g_cache_order.pop_front();
        }
    }
    g_fp_cache.emplace(fp, ins.first->second);
    return ins.first->second;
}

}  // namespace qf
