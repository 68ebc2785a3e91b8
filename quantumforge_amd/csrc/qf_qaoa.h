// QAOA-form instances of an uneven batch (qf_uneven_adjoint_gradient,
// qf_uneven_expectation) on a tile engine of their own.
#pragma once
#include <cstdint>
#include <vector>

#include "qf_runtime.h"

namespace qfq {

// A circuit in QAOA form: a Hadamard on every qubit, then layers of cost
// rotations exp(-i theta/2 Z_u Z_v) (written CNOT(u, v) Rz(theta)_v CNOT(u, v))
// followed by an X rotation on every qubit; every rotation is a parameter
// occurrence (slot = its rank among the circuit's rotations).  The observable
// is a real sum of Z and Z Z words and a constant.
struct Form {
    struct Edge {
        uint32_t u, v, slot;
    };
    struct Layer {
        std::vector<Edge> edges;
        std::vector<uint32_t> mix;  // RX slot of every qubit
    };
    struct Term {
        uint32_t a, b;  // qubits (b == a: a single Z; a == b == ~0u: never stored, see c0)
        double coef;
    };
    uint32_t n = 0;
    std::vector<Layer> layers;
    std::vector<Term> obs;
    double c0 = 0.0;
};

// True when instance `q` (already validated; pc built by validate_ops) is in
// QAOA form with 12 <= n <= 32 and at most kMaxTerms diagonal terms per layer
// and in the observable.
bool match(const qf_uneven_instance& q, const qfrt::PCirc& pc, Form& out);

// Energies and (want_grad) occurrence gradients of the instances ids[] (all in
// QAOA form): grads of instance k start at grad_off[k].  Adds to the stats'
// wave, graph and byte counters.
qf_status run(qf_backend* be, const qf_uneven_instance* inst, const std::vector<uint32_t>& ids,
              const std::vector<qfrt::PCirc>& pcs, const std::vector<Form>& forms,
              const std::vector<uint64_t>& grad_off, int32_t precision, bool want_grad, double* energies,
              double* grads, qf_uneven_stats& S);

}  // namespace qfq
