// Host verifier entry (see verifier.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace gkrx {

// Gate lists of a circuit with canonical coefficients; layer 0 is the input layer.
struct HostCircuitView {
    uint32_t D;
    const uint32_t *n, *m;                          // nb_input_vars / nb_output_vars
    const std::vector<std::vector<uint32_t>> &mul;  // (in0, in1, out, coef)
    const std::vector<std::vector<uint32_t>> &add;  // (in0, out, coef)
};

// 1 = accepted, 0 = rejected.
int host_verify(const HostCircuitView &c, uint32_t L, uint32_t R, int kind, uint32_t grinding_bits,
                const uint32_t *claimed, const uint8_t *proof, size_t proof_len);
// The same over the BN254 scalar field (one lane, one repetition); claimed is the
// canonical value as 8 little-endian u32 words.
int host_verify_bn254(const HostCircuitView &c, int kind, uint32_t grinding_bits, const uint32_t *claimed,
                      const uint8_t *proof, size_t proof_len);

}  // namespace gkrx
