// The HeOps surface (crates/ckks-train-ckks/src/ops.rs:198-356) over the device engine, with
// the checks, errors and counts of the CPU Evaluator (crates/ckks-train-ckks/src/evaluator.rs).
#pragma once

#include <functional>
#include <map>
#include <mutex>

#include "runtime.h"

namespace ckh {

// Polynomial-level operations used by the evaluator and exported through the C ABI.
PolyPtr poly_alloc(Context& c, u32 level, u32 p_limbs, bool ntt);
PolyPtr poly_upload(Context& c, const u64* host, u32 level, u32 p_limbs, bool ntt);
void poly_download(Context& c, const Poly& p, u64* host);
u64 poly_digest(Context& c, const Poly& p);
PolyPtr poly_uniform(Context& c, u32 level, u32 p_limbs, bool ntt, u64 seed, u64 stream0);
PolyPtr poly_transform(Context& c, const Poly& in, bool forward);
PolyPtr poly_automorphism(Context& c, const Poly& in, u64 g);
PolyPtr poly_rescale(Context& c, const Poly& in);
double time_ntt(Context& c, const std::vector<const Poly*>& polys, bool forward, u32 reps);
// Mean ms of a streaming device copy of bytes (two fresh buffers), reps launches on the stream.
double time_copy(Context& c, u64 bytes, u32 reps);
PolyPtr poly_mod_down(Context& c, const Poly& qp);
PolyPtr poly_pointwise(Context& c, PwOp op, const Poly& a, const Poly* b);
PolyPtr poly_scalar_mul(Context& c, const Poly& a, u64 k, bool neg);
PolyPtr poly_raise(Context& c, const Poly& a, u32 top);
// keyswitch::decompose: one buffer holding all digits; returns one QpPoly handle per digit.
std::vector<PolyPtr> ks_decompose(Context& c, const Poly& d);
// keyswitch::apply_key for digits laid out as ks_decompose returns them (or copied so).
std::pair<PolyPtr, PolyPtr> ks_apply(Context& c, const std::vector<PolyPtr>& digits, const Ksk& key);
std::pair<PolyPtr, PolyPtr> ks_switch(Context& c, const Poly& d, const Ksk& key);
// n switches under one key in batched launches.
std::vector<std::pair<PolyPtr, PolyPtr>> ks_switch_batch(Context& c, const std::vector<const Poly*>& d,
                                                         const Ksk& key);
// The same switches into caller-given outputs, with no device allocation: intermediates and
// descriptors in the GPU's fixed workspace (Context::static_scratch), launched without a graph.
void ks_switch_batch_static(Context& c, const std::vector<const Poly*>& d, const Ksk& key,
                            const std::vector<Poly*>& kb, const std::vector<Poly*>& ka);
KskPtr ksk_upload(Context& c, const u64* host, u32 max_level, u64 galois);
KskPtr ksk_uniform(Context& c, u32 max_level, u64 galois, u64 seed);

u64 galois_elt(u64 n, int32_t step);

// EvalConfig (crates/ckks-train-ckks/src/ops.rs:111-134): the optimised (all true, the default)
// or naive form of hoisted rotations (one shared decomposition), BSGS products (one giant
// rotation per block) and relinearization of product sums (once per sum).
struct EvalConfig {
    bool hoist = true, bsgs = true, lazy_relin = true;
};

// A BSGS product's shape: entries[k1 * baby + k2] hold rot(diag_{offset + baby k1 + k2},
// -(offset + baby k1) stride).
struct BsgsShape {
    u32 baby, giant, stride;
    int32_t offset;
};

// One stage of CoeffToSlot or SlotToCoeff (LinearTransform,
// crates/ckks-train-ckks/src/bootstrap/linear.rs:467-502): plaintext diagonals at the stage's
// level in a BSGS shape (null entries are zero and skipped), the diagonals of a two-input stage
// (CoeffToSlot's last, SlotToCoeff's first), and the exact scale after the stage's rescale.
struct LinearStage {
    bool slot_to_coeff;  // the direction of the stage list it was built for
    u32 level;           // level of the stage input and of its diagonals
    double scale;        // exact scale after the stage's rescale (at level - 1)
    BsgsShape shape;
    std::vector<const Poly*> diags;  // baby * giant entries
    double pt_scale;                 // the common scale of the diagonals
    bool has_second;
    BsgsShape second_shape;
    std::vector<const Poly*> second;
    double second_pt_scale;
};

class Evaluator {
   public:
    Evaluator(Context& c, KskPtr relin, std::vector<KskPtr> galois);
    Context& ctx;
    KskPtr relin;
    std::map<u64, KskPtr> galois;
    EvalConfig config;

    ckh_op_counter counter();
    void reset_counter();
    void add_counts(const ckh_op_counter& c);
    bool has_galois_key(int32_t step) const;
    bool owns(const Ct& ct) const;

    CtPtr add(const Ct& a, const Ct& b, bool sub);
    CtPtr negate(const Ct& a);
    CtPtr add_plain(const Ct& a, const Poly& pt, double pt_scale);
    CtPtr mul_plain(const Ct& a, const Poly& pt, double pt_scale);
    CtPtr mul_no_relin(const Ct& a, const Ct& b);
    // mul_no_relin of several pairs in one launch (the server's turns).
    std::vector<CtPtr> mul_no_relin_batch(const std::vector<std::pair<const Ct*, const Ct*>>& pairs);
    CtPtr relinearize(const Ct& ct);
    std::vector<CtPtr> relinearize_batch(const std::vector<const Ct*>& cts);
    CtPtr mul(const Ct& a, const Ct& b);
    CtPtr rescale(const Ct& ct);
    // rescale of several ciphertexts, at any levels, in one program (the server's turns).
    std::vector<CtPtr> rescale_batch(const std::vector<const Ct*>& cts);
    CtPtr mod_switch_to_level(const Ct& ct, u32 level);
    CtPtr relabel(const Ct& ct, double scale);
    CtPtr rotate(const Ct& ct, int32_t steps);
    CtPtr conjugate(const Ct& ct);
    std::vector<CtPtr> hoisted_rotate(const Ct& ct, const std::vector<int32_t>& steps);
    // hoisted_rotate of several ciphertexts, each with its own steps, in batched launches.
    std::vector<std::vector<CtPtr>> hoisted_rotate_batch(
        const std::vector<std::pair<const Ct*, std::vector<int32_t>>>& items);
    // Rotations of several ciphertexts (their own steps) switched in one batch.
    std::vector<CtPtr> rotate_batch(const std::vector<std::pair<const Ct*, int32_t>>& items);
    CtPtr matvec_plain(const Ct& x, const std::vector<const Poly*>& entries, double pt_scale, u32 baby, u32 giant,
                       u32 stride, int32_t offset);
    CtPtr matvec_enc(const Ct& x, const std::vector<const Ct*>& entries, u32 baby, u32 giant, u32 stride,
                     int32_t offset);
    CtPtr matvec_plain_sum(const std::vector<const Ct*>& xs, const std::vector<std::vector<const Poly*>>& entries,
                           double pt_scale, u32 baby, u32 giant, u32 stride, int32_t offset);
    std::vector<CtPtr> matvec_enc_batch(const std::vector<const Ct*>& xs,
                                        const std::vector<std::vector<const Ct*>>& entries,
                                        const std::vector<BsgsShape>& shapes);
    u64 matvec_enc_workspace_bytes(u32 level, u32 baby, u32 giant) const;
    CtPtr mul_scalar(const Ct& ct, double value, double target_scale);
    CtPtr mul_plain_to_scale(const Ct& a, const Poly& pt, double pt_scale, double target_scale);
    double exact_product_scale(u32 level, double target_scale) const;
    double target_plain_scale(const Ct& ct, double target_scale) const;
    CtPtr mod_raise(const Ct& ct, const Ksk& to_sparse, const Ksk& to_dense);
    CtPtr mod_raise_main_key(const Ct& ct);
    CtPtr mul_lazy(const Ct& a, const Ct& b);
    CtPtr relin_lazy(const Ct& sum);
    // CoeffToSlot / SlotToCoeff (linear.rs:776-859) over the BSGS products.
    std::pair<CtPtr, CtPtr> coeff_to_slot(const Ct& ct, const std::vector<LinearStage>& stages);
    CtPtr slot_to_coeff(const Ct& re, const Ct& im, const std::vector<LinearStage>& stages);

   private:
    std::mutex mu_;
    ckh_op_counter cnt_{};
    void count(u64 ckh_op_counter::*field, u64 v = 1);
    // Runs a composite operation whose counts are merged only if it succeeds.
    template <class F>
    auto counted(F&& body) -> decltype(body());
    void check_basis(const Ct& ct) const;
    void check_mul_operands(const Ct& a, const Ct& b) const;
    const Ksk& relin_key_checked(u32 level) const;
    const Ksk& galois_key(u64 elt, u32 level) const;
    u64 rotation_elt(int32_t steps) const;  // 0 for a multiple of the slot count
    CtPtr apply_galois(const Ct& ct, u64 elt);
    void check_mod_raise_input(const Ct& ct) const;
    CtPtr raise_ct(const Ct& ct);
    std::vector<bool> check_matvec_enc(const Ct& x, const std::vector<const Ct*>& entries, u32 baby, u32 giant,
                                       u32 stride, int32_t offset) const;
    CtPtr matvec_enc_run(const Ct& x, const std::vector<const Ct*>& entries, u32 baby, u32 giant, u32 stride,
                         int32_t offset, const std::vector<bool>& present);
    // The BSGS shape checks shared by both products, before any work.
    void check_bsgs_shape(u32 level, u32 baby, u32 giant, u32 stride, int32_t offset,
                          const std::vector<bool>& present) const;
    // The BSGS driver (linalg.rs:800-888): for every term t, the present[k1 * baby + k2]
    // entries are multiplied by rot(x_t, k2 stride) with mul(t, index, rotated x_t) and each
    // block (or, naive, each product) is finished with finish() before its giant rotation.
    struct BsgsTerm {
        const Ct* x;
        std::vector<bool> present;
    };
    CtPtr sum_cts(const std::vector<CtPtr>& cts);
    void check_stages(const std::vector<LinearStage>& stages, u32 level, double scale) const;
    CtPtr apply_stage(const LinearStage& lt, const Ct& x, const Ct* x2);
    CtPtr mul_by_x_half(const Ct& ct);
    std::once_flag mono_once_;
    PolyPtr mono_;  // X^(N/2) in NTT form over every Q prime, built at first use
    std::vector<std::vector<CtPtr>> baby_rotations(const std::vector<BsgsTerm>& terms, u32 baby, u32 giant,
                                                   u32 stride);
    CtPtr plain_bsgs(const std::vector<BsgsTerm>& terms, const std::vector<std::vector<const Poly*>>& entries,
                     double pt_scale, u32 baby, u32 giant, u32 stride, int32_t offset);
    CtPtr bsgs(const std::vector<BsgsTerm>& terms, u32 baby, u32 giant, u32 stride, int32_t offset,
               const std::function<CtPtr(size_t, size_t, const Ct&)>& mul,
               const std::function<CtPtr(const CtPtr&)>& finish);
};

}  // namespace ckh
