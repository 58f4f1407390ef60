// Drop-in C++ surface of the reference prover (PolyhedraZK/Expander-cpp) over the gkrx
// C ABI.  Call sites written against the reference --
//
//     Config config;
//     auto circuit = gkr::Circuit<F, F_primitive>::load_extracted_gates(mul, add);
//     circuit.layers[0].input_layer_vals.evals = witness;      // or set_random_input()
//     circuit.evaluate();
//     gkr::Prover<F, F_primitive> prover(config);
//     prover.prepare_mem(circuit);
//     auto [claimed_v, proof] = prover.prove(circuit);
//     gkr::Verifier verifier(config);
//     bool ok = verifier.verify(circuit, claimed_v, proof);
//
// -- compile unchanged against this header (include/LinearGKR/LinearGKR.hpp:34-150,
// include/circuit/circuit.hpp:13-305, include/configuration/config.hpp:5-86,
// include/fiat_shamir/transcript.hpp:13-85, include/field/M31.hpp:35-449).  The prove
// runs on the GPU; the proof bytes are the reference's.
//
// F = gkr::M31_field::VectorizedM31 has GKRX_M31_LANES lanes (8, as the x86 reference
// ships; define GKRX_M31_LANES=16 before including for 16).  Errors the reference turns
// into asserts abort here too; the verifier returns false on rejection.
#pragma once

#include <algorithm>
#include <cassert>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../gkrx.h"

#ifndef GKRX_M31_LANES
#define GKRX_M31_LANES 8
#endif

// config.hpp:5-24 (global enums in the reference)
enum Polynomial_commitment_type { Raw, KZG, Orion, FRI };
enum Field_type { M31, BabyBear, BN254 };
enum FiatShamir_hash_type { SHA256, Keccak256, Poseidon, Animoe, MIMC7 };

// config.hpp:26-86
class Config {
    int num_repetitions;

  public:
    int field_size;
    int security_bits;
    int grinding_bits;
    int nb_parallel;
    Polynomial_commitment_type PC_type;
    Field_type field_type;
    FiatShamir_hash_type FS_hash;

    void initialize_config()
    {
        security_bits = 100;
        field_size = field_type == BN254 ? 254 : 31;
        num_repetitions = (security_bits - grinding_bits) / field_size;
        if ((security_bits - grinding_bits) % field_size != 0) num_repetitions++;
    }
    Config()
    {
        security_bits = 100;
        grinding_bits = 10;
        nb_parallel = 16;
        PC_type = Raw;
        field_type = M31;
        FS_hash = SHA256;
        initialize_config();
    }
    int get_num_repetitions() const { return num_repetitions; }
};

namespace gkr {

typedef unsigned char uint8;
typedef unsigned uint32;
typedef unsigned long long uint64;

namespace M31_field {

const int mod = 2147483647;

// M31.hpp:35-142: canonical u32 scalar
class M31 {
  public:
    uint32 x;
    M31() : x(0) {}
    M31(uint32 v) { x = (v & (uint32)mod) + (v >> 31); }
    static M31 new_unchecked(uint32 v) { M31 f; f.x = v; return f; }
    static M31 zero() { return new_unchecked(0); }
    static M31 one() { return new_unchecked(1); }
    M31 operator+(const M31 &r) const { uint32 s = x + r.x; return new_unchecked(s >= (uint32)mod ? s - mod : s); }
    M31 operator-() const { return new_unchecked(x == 0 ? 0 : mod - x); }
    M31 operator-(const M31 &r) const { return *this + (-r); }
    M31 operator*(const M31 &r) const
    {
        uint64 p = (uint64)x * r.x;
        p = (p & (uint64)mod) + (p >> 31);
        if (p >= (uint64)mod) p -= mod;
        return new_unchecked((uint32)p);
    }
    void operator+=(const M31 &r) { *this = *this + r; }
    void operator*=(const M31 &r) { *this = *this * r; }
    bool operator==(const M31 &r) const { return x == r.x; }
    bool operator!=(const M31 &r) const { return x != r.x; }
    void to_bytes(uint8 *out) const { memcpy(out, &x, 4); }
    void from_bytes(const uint8 *in)
    {
        memcpy(&x, in, 4);
        x = (x & (uint32)mod) + (x >> 31);
        if (x >= (uint32)mod) x -= mod;
    }
    static int byte_length() { return 4; }
    static M31 random() { return M31((uint32)rand()); }
};

// M31.hpp:240-449: GKRX_M31_LANES data-parallel lanes, lane 0 first in memory
template <int LANES>
class VectorizedM31T {
  public:
    typedef M31 primitive_type;
    uint32 lanes[LANES];
    VectorizedM31T() { memset(lanes, 0, sizeof lanes); }
    VectorizedM31T(uint32 v)
    {
        M31 f(v);
        for (int i = 0; i < LANES; i++) lanes[i] = f.x;
    }
    static VectorizedM31T zero() { return VectorizedM31T(); }
    static VectorizedM31T one() { return VectorizedM31T(1u); }
    static size_t pack_size() { return LANES; }
    static VectorizedM31T random()
    {
        VectorizedM31T r;
        for (int i = 0; i < LANES; i++) {
            const uint32 v = (uint32)rand();  // PackedM31::random: rand() reduced to [0, p)
            r.lanes[i] = v >= (uint32)mod ? v - mod : v;
        }
        return r;
    }
    static VectorizedM31T random_bool()
    {
        VectorizedM31T r;
        for (int i = 0; i < LANES; i++) r.lanes[i] = (uint32)(rand() % 2);
        return r;
    }
    bool operator==(const VectorizedM31T &o) const { return memcmp(lanes, o.lanes, sizeof lanes) == 0; }
    bool operator!=(const VectorizedM31T &o) const { return !(*this == o); }
    void to_bytes(uint8 *out) const { memcpy(out, lanes, sizeof lanes); }
    void from_bytes(const uint8 *in)
    {
        for (int i = 0; i < LANES; i++) {
            M31 f;
            f.from_bytes(in + 4 * i);
            lanes[i] = f.x;
        }
    }
    std::vector<M31> unpack() const
    {
        std::vector<M31> r;
        for (int i = 0; i < LANES; i++) r.push_back(M31::new_unchecked(lanes[i]));
        return r;
    }
};
typedef VectorizedM31T<GKRX_M31_LANES> VectorizedM31;

}  // namespace M31_field

// circuit.hpp:13-30
template <typename F, uint32 nb_input>
class Gate {
  public:
    uint32 o_id;
    uint32 i_ids[nb_input];
    F coef;
};

// circuit.hpp:32-66
template <typename F, uint32 nb_input>
class SparseCircuitConnection {
  public:
    uint32 nb_output_vars = 0, nb_input_vars = 0;
    std::vector<Gate<F, nb_input>> sparse_evals;
};

// poly.hpp:106-124
template <typename F>
class MultiLinearPoly {
  public:
    uint32 nb_vars = 0;
    std::vector<F> evals;
};

// circuit.hpp:68-116
template <typename F, typename F_primitive>
class CircuitLayer {
  public:
    uint32 nb_output_vars = 0, nb_input_vars = 0;
    MultiLinearPoly<F> input_layer_vals, output_layer_vals;
    SparseCircuitConnection<F_primitive, 1> add;
    SparseCircuitConnection<F_primitive, 2> mul;
    uint32 nb_mul_gates() const { return (uint32)mul.sparse_evals.size(); }
    uint32 nb_add_gates() const { return (uint32)add.sparse_evals.size(); }
};

namespace detail {

inline void check(int rc, const char *what)
{
    if (rc < 0) {
        fprintf(stderr, "gkrx: %s failed (%d): %s\n", what, rc, gkrx_last_error());
        abort();
    }
}

inline gkrx_config to_c(const Config &cfg, int lanes)
{
    gkrx_config c;
    gkrx_config_default(&c);
    c.field_type = cfg.field_type == BN254 ? GKRX_FIELD_BN254 : (cfg.field_type == BabyBear ? GKRX_FIELD_BABYBEAR : GKRX_FIELD_M31);
    c.fs_hash = (int)cfg.FS_hash;
    c.pc_type = (int)cfg.PC_type;
    c.security_bits = cfg.security_bits;
    c.grinding_bits = cfg.grinding_bits;
    c.nb_parallel = lanes;
    return c;
}

// One prover context per (config, lane count) in the process, created on first use.
inline gkrx_ctx *context_for(const Config &cfg, int lanes)
{
    struct Entry { gkrx_config c; gkrx_ctx *ctx; };
    static std::vector<Entry> cache;
    gkrx_config c = to_c(cfg, lanes);
    for (auto &e : cache)
        if (memcmp(&e.c, &c, sizeof c) == 0) return e.ctx;
    gkrx_ctx *ctx = nullptr;
    check(gkrx_ctx_create(&c, 0, &ctx), "gkrx_ctx_create");
    cache.push_back({c, ctx});
    return ctx;
}

}  // namespace detail

// circuit.hpp:118-305.  The device copy of the wiring is built on first use.
template <typename F, typename F_primitive>
class Circuit {
  public:
    std::vector<CircuitLayer<F, F_primitive>> layers;

    Circuit() = default;
    Circuit(const Circuit &o) : layers(o.layers) {}
    Circuit &operator=(const Circuit &o)
    {
        layers = o.layers;
        dev_.reset();
        return *this;
    }

    static Circuit load_extracted_gates(const char *filename_mul, const char *filename_add)
    {
        Circuit c;
        FILE *f = fopen(filename_mul, "r");
        if (!f) { fprintf(stderr, "cannot open %s\n", filename_mul); abort(); }
        uint32 nb = 0;
        int ch, last = '\n';
        while ((ch = fgetc(f)) != EOF) { if (ch == '\n') nb++; last = ch; }
        if (last != '\n') nb++;
        rewind(f);
        c.layers.resize(nb);
        for (uint32 i = 0; i < nb; i++) {
            auto &layer = c.layers[nb - 1 - i];
            unsigned long ng = 0;
            if (fscanf(f, "%lu", &ng) != 1) abort();
            layer.mul.sparse_evals.resize(ng);
            for (unsigned long g = 0; g < ng; g++) {
                unsigned long i0, i1, o;
                unsigned coef;
                if (fscanf(f, "%lu%lu%lu%u", &i0, &i1, &o, &coef) != 4) abort();
                auto &gt = layer.mul.sparse_evals[g];
                gt.o_id = (uint32)o; gt.i_ids[0] = (uint32)i0; gt.i_ids[1] = (uint32)i1; gt.coef = F_primitive(coef);
            }
        }
        fclose(f);
        f = fopen(filename_add, "r");
        if (!f) { fprintf(stderr, "cannot open %s\n", filename_add); abort(); }
        for (uint32 i = 0; i < nb; i++) {
            auto &layer = c.layers[nb - 1 - i];
            unsigned long ng = 0;
            if (fscanf(f, "%lu", &ng) != 1) abort();
            layer.add.sparse_evals.resize(ng);
            for (unsigned long g = 0; g < ng; g++) {
                unsigned long i0, o;
                unsigned coef;
                if (fscanf(f, "%lu%lu%u", &i0, &o, &coef) != 3) abort();
                auto &gt = layer.add.sparse_evals[g];
                gt.o_id = (uint32)o; gt.i_ids[0] = (uint32)i0; gt.coef = F_primitive(coef);
            }
        }
        fclose(f);
        c._compute_nb_vars();
        return c;
    }

    // circuit.hpp:124-149
    void _compute_nb_vars()
    {
        auto vars = [](uint32 id) -> uint32 {
            if (id == 0) return 0;
            uint32 p = (id & (id - 1)) == 0 ? id : (1u << (32 - __builtin_clz(id)));
            return (uint32)__builtin_ctz(p);
        };
        for (auto &layer : layers) {
            uint32 mo = 0, mi = 0;
            for (auto &g : layer.mul.sparse_evals) {
                mo = std::max(mo, g.o_id);
                mi = std::max(mi, std::max(g.i_ids[0], g.i_ids[1]));
            }
            for (auto &g : layer.add.sparse_evals) {
                mo = std::max(mo, g.o_id);
                mi = std::max(mi, g.i_ids[0]);
            }
            layer.nb_input_vars = vars(mi);
            layer.nb_output_vars = vars(mo);
            layer.input_layer_vals.nb_vars = layer.nb_input_vars;
        }
    }

    uint32 log_input_size() const { return layers[0].nb_input_vars; }

    // circuit.hpp:287-304 (rand()-based, like the reference)
    void set_random_input()
    {
        auto &v = layers[0].input_layer_vals.evals;
        v.clear();
        for (uint32 i = 0; i < (1u << log_input_size()); i++) v.emplace_back(F::random());
    }
    void set_random_boolean_input()
    {
        auto &v = layers[0].input_layer_vals.evals;
        v.clear();
        for (uint32 i = 0; i < (1u << log_input_size()); i++) v.emplace_back(F::random_bool());
    }

    uint32 nb_mul_gates() const { uint32 s = 0; for (auto &l : layers) s += l.nb_mul_gates(); return s; }
    uint32 nb_add_gates() const { uint32 s = 0; for (auto &l : layers) s += l.nb_add_gates(); return s; }

    // circuit.hpp:273-280: every layer's values are computed on the device and copied
    // back into input_layer_vals / output_layer_vals.
    void evaluate()
    {
        gkrx_circuit *d = device(detail::context_for(Config(), (int)F::pack_size()));
        upload_witness(d);
        detail::check(gkrx_circuit_evaluate(d), "gkrx_circuit_evaluate");
        for (size_t l = 0; l <= layers.size(); l++) {
            const uint32 vars = l < layers.size() ? layers[l].nb_input_vars : layers.back().nb_output_vars;
            std::vector<F> vals((size_t)1 << vars);
            detail::check(gkrx_circuit_get_layer_values(d, (uint32_t)l, reinterpret_cast<uint32_t *>(vals.data()),
                                                        vals.size() * F::pack_size()),
                          "gkrx_circuit_get_layer_values");
            if (l == 0) continue;
            if (l < layers.size()) layers[l].input_layer_vals.evals = std::move(vals);
            else layers.back().output_layer_vals.evals = std::move(vals);
        }
    }

    // the device circuit for a given context (built once per context)
    gkrx_circuit *device(gkrx_ctx *ctx) const
    {
        if (dev_ && dev_->ctx == ctx) return dev_->c;
        std::vector<std::vector<uint32_t>> mul(layers.size()), add(layers.size());
        std::vector<uint32_t> nm(layers.size()), na(layers.size());
        std::vector<const uint32_t *> mp(layers.size()), ap(layers.size());
        for (size_t l = 0; l < layers.size(); l++) {
            for (auto &g : layers[l].mul.sparse_evals) mul[l].insert(mul[l].end(), {g.i_ids[0], g.i_ids[1], g.o_id, g.coef.x});
            for (auto &g : layers[l].add.sparse_evals) add[l].insert(add[l].end(), {g.i_ids[0], g.o_id, g.coef.x});
            nm[l] = (uint32_t)layers[l].mul.sparse_evals.size();
            na[l] = (uint32_t)layers[l].add.sparse_evals.size();
            mp[l] = mul[l].data();
            ap[l] = add[l].data();
        }
        gkrx_circuit *c = nullptr;
        detail::check(gkrx_circuit_from_arrays(ctx, (uint32_t)layers.size(), nm.data(), mp.data(), na.data(),
                                               ap.data(), &c),
                      "gkrx_circuit_from_arrays");
        dev_ = std::make_shared<Dev>(ctx, c);
        return c;
    }

    void upload_witness(gkrx_circuit *d) const
    {
        const auto &w = layers[0].input_layer_vals.evals;
        detail::check(gkrx_circuit_set_witness(d, reinterpret_cast<const uint32_t *>(w.data()), w.size() * F::pack_size()),
                      "gkrx_circuit_set_witness");
    }

  private:
    struct Dev {
        gkrx_ctx *ctx;
        gkrx_circuit *c;
        Dev(gkrx_ctx *x, gkrx_circuit *y) : ctx(x), c(y) {}
        ~Dev() { gkrx_circuit_destroy(c); }
    };
    mutable std::shared_ptr<Dev> dev_;
};

// transcript.hpp:13-85
template <typename F>
class Proof {
  public:
    uint32 idx = 0, commitment_nb_bytes = 0, opening_nb_bytes = 0;
    std::vector<uint8> bytes;
    void append_bytes(const uint8 *in, uint32 len) { bytes.insert(bytes.end(), in, in + len); }
    void reset() { idx = 0; }
    const uint8 *bytes_head() { return bytes.data() + idx; }
    void step(uint32 n) { idx += n; }
    F get_next_and_step()
    {
        F f;
        f.from_bytes(bytes.data() + idx);
        step(sizeof(F));
        return f;
    }
    // transcript.hpp:54-85: [u32 byte count][bytes]; the compressed copy is made with tar as
    // the reference does.  Both paths reach the shell single-quoted (a quote inside one is
    // written '\''), so a path cannot add words or commands to the tar line.
    void write_and_compress(const char *proof_loc = "/tmp/proof.txt",
                            const char *compressed_proof_loc = "/tmp/compressed_proof.tar.gz")
    {
        if (gkrx_proof_write(proof_loc, bytes.data(), bytes.size()) != GKRX_OK) std::abort();
        if (compressed_proof_loc) {
            const std::string cmd = std::string("tar -czf ") + shell_quote(compressed_proof_loc) +
                                    " --absolute-names -- " + shell_quote(proof_loc);
            if (std::system(cmd.c_str()) != 0) std::abort();
        }
    }
    static std::string shell_quote(const char *s)
    {
        std::string q = "'";
        for (; *s; ++s) {
            if (*s == '\'') q += "'\\''";
            else q += *s;
        }
        return q + "'";
    }
    void read_proof(const char *proof_loc = "/tmp/proof.txt")
    {
        size_t n = 0;
        if (gkrx_proof_read(proof_loc, nullptr, 0, &n) != GKRX_OK) std::abort();
        bytes.resize(n);
        if (gkrx_proof_read(proof_loc, bytes.data(), bytes.size(), &n) != GKRX_OK) std::abort();
        idx = 0;
    }
};

// LinearGKR.hpp:34-91
template <typename F, typename F_primitive>
class Prover {
  public:
    const Config &config;
    explicit Prover(const Config &config_) : config(config_)
    {
        assert(config.field_type == Field_type::M31);
        assert(config.PC_type == Polynomial_commitment_type::Raw);
    }
    // Device memory is owned by the context; kept so that call sites pair it with prove.
    void prepare_mem(const Circuit<F, F_primitive> &) {}

    std::tuple<std::vector<F>, Proof<F>> prove(const Circuit<F, F_primitive> &circuit)
    {
        gkrx_ctx *ctx = detail::context_for(config, (int)F::pack_size());
        gkrx_circuit *d = circuit.device(ctx);
        circuit.upload_witness(d);
        Proof<F> proof;
        proof.bytes.resize(gkrx_proof_size(ctx, d));
        std::vector<F> claimed((size_t)config.get_num_repetitions());
        size_t len = 0;
        detail::check(gkrx_prove(ctx, d, proof.bytes.data(), proof.bytes.size(), &len,
                                 reinterpret_cast<uint32_t *>(claimed.data())),
                      "gkrx_prove");
        proof.bytes.resize(len);
        return {claimed, proof};
    }
};

// LinearGKR.hpp:93-150
class Verifier {
  public:
    const Config &config;
    explicit Verifier(const Config &config_) : config(config_) {}

    template <typename F, typename F_primitive>
    bool verify(const Circuit<F, F_primitive> &circuit, const std::vector<F> &claimed_v, Proof<F> &proof)
    {
        gkrx_ctx *ctx = detail::context_for(config, (int)F::pack_size());
        gkrx_circuit *d = circuit.device(ctx);
        int rc = gkrx_verify(ctx, d, reinterpret_cast<const uint32_t *>(claimed_v.data()), proof.bytes.data() + proof.idx,
                             proof.bytes.size() - proof.idx);
        detail::check(rc, "gkrx_verify");
        proof.idx = (uint32)proof.bytes.size();
        return rc == 1;
    }
};

}  // namespace gkr
