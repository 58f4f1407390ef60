// Reference-style call site compiled against include/gkrx/expander.hpp: it is written
// exactly as a user of the reference's C++ API would write it (LinearGKR.hpp:34-150).
//   prove_example <mul.txt> <add.txt> <witness.bin> [expected_proof.bin] [keccak]
// Exit 0 when the verifier accepts (and the proof equals the expected bytes, if given).
#include <cstdio>
#include <vector>

#include "gkrx/expander.hpp"

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s mul.txt add.txt witness.bin [expected_proof.bin] [keccak]\n", argv[0]);
        return 2;
    }
    using F = gkr::M31_field::VectorizedM31;
    using F_primitive = gkr::M31_field::M31;
    Config config;
    if (argc > 5 && std::string(argv[5]) == "keccak") config.FS_hash = Keccak256;
    auto circuit = gkr::Circuit<F, F_primitive>::load_extracted_gates(argv[1], argv[2]);
    FILE *f = fopen(argv[3], "rb");
    if (!f) return 2;
    auto &evals = circuit.layers[0].input_layer_vals.evals;
    evals.resize((size_t)1 << circuit.log_input_size());
    for (auto &e : evals) {
        unsigned char buf[sizeof(F)];
        if (fread(buf, 1, sizeof buf, f) != sizeof buf) return 2;
        e.from_bytes(buf);
    }
    fclose(f);
    circuit.evaluate();

    gkr::Prover<F, F_primitive> prover(config);
    prover.prepare_mem(circuit);
    auto t = prover.prove(circuit);
    auto claimed_v = std::get<0>(t);
    auto proof = std::get<1>(t);

    bool same = true;
    if (argc > 4 && std::string(argv[4]) != "-") {
        FILE *g = fopen(argv[4], "rb");
        std::vector<unsigned char> want;
        int c;
        while (g && (c = fgetc(g)) != EOF) want.push_back((unsigned char)c);
        if (g) fclose(g);
        same = want == proof.bytes;
    }
    gkr::Verifier verifier(config);
    bool ok = verifier.verify(circuit, claimed_v, proof);
    printf("proof %zu bytes, %s, verifier %s\n", proof.bytes.size(), same ? "matches expected" : "DIFFERS",
           ok ? "accepted" : "rejected");
    return ok && same ? 0 : 1;
}
