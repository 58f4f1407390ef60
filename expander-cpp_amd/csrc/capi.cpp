// C ABI of the prover (include/gkrx.h): contexts, circuits, prove, verify.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "bn254.hpp"
#include "gkrx.h"
#include "host_hash.hpp"
#include "internal.hpp"
#include "handles.hpp"
#include "verifier.hpp"

namespace gkrx {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

}  // namespace gkrx

using namespace gkrx;

extern "C" const char *gkrx_last_error(void) { return g_err.c_str(); }

extern "C" void gkrx_config_default(gkrx_config *cfg)
{
    if (!cfg) return;
    cfg->field_type = GKRX_FIELD_M31;
    cfg->fs_hash = GKRX_FS_SHA256;
    cfg->pc_type = GKRX_PC_RAW;
    cfg->security_bits = 100;
    cfg->grinding_bits = 10;
    cfg->nb_parallel = 8;
}

static int field_bits(int field_type)
{
    switch (field_type) {
    case GKRX_FIELD_M31: return 31;
    case GKRX_FIELD_BABYBEAR: return 31;
    case GKRX_FIELD_BN254: return 254;
    default: return -1;
    }
}

extern "C" int gkrx_config_num_repetitions(const gkrx_config *cfg)
{
    if (!cfg) return GKRX_ERR_ARG;
    const int fs = field_bits(cfg->field_type);
    if (fs <= 0) return GKRX_ERR_CONFIG;
    const int num = cfg->security_bits - cfg->grinding_bits;
    return num / fs + (num % fs != 0 ? 1 : 0);
}

extern "C" int gkrx_ctx_create(const gkrx_config *cfg, int device, gkrx_ctx **out)
{
    if (!cfg || !out) return GKRX_ERR_ARG;
    // Prover::Prover asserts M31 / Raw (LinearGKR.hpp:44-46); both transcript hashes run
    // here, and so does the BN254 scalar field (one lane) that the Config enum names.
    const bool bn = cfg->field_type == GKRX_FIELD_BN254;
    if ((cfg->field_type != GKRX_FIELD_M31 && !bn) || cfg->pc_type != GKRX_PC_RAW ||
        (cfg->fs_hash != GKRX_FS_SHA256 && cfg->fs_hash != GKRX_FS_KECCAK256) ||
        (!bn && cfg->nb_parallel != 8 && cfg->nb_parallel != 16) || cfg->grinding_bits < 0 ||
        cfg->grinding_bits > 24) {
        set_error("unsupported config");
        return GKRX_ERR_CONFIG;
    }
    const int R = gkrx_config_num_repetitions(cfg);
    if (R < 1 || R > 3) {
        set_error("num_repetitions must be 1..3 (the reference keeps 3 helpers, sumcheck.hpp:65)");
        return GKRX_ERR_CONFIG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        set_error("no usable HIP device");
        return GKRX_ERR_DEVICE;
    }
    GKRX_HIP_CHECK(hipSetDevice(device));
    gkrx_ctx *c = new (std::nothrow) gkrx_ctx;
    if (!c) return GKRX_ERR_NOMEM;
    c->cfg = *cfg;
    c->device = device;
    c->R = (uint32_t)R;
    c->bn254 = bn;
    c->L = bn ? 1u : (uint32_t)cfg->nb_parallel;
    c->W = bn ? 8u : c->L;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return GKRX_ERR_DEVICE;
    }
    // the allocation-free prover's table array on this device (gkrx_prove_noalloc)
    if (int rc = dev_noalloc_init(device)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_destroy(gkrx_ctx *ctx)
{
    if (!ctx) return GKRX_ERR_ARG;
    if (ctx->live_circuits.load() != 0) {
        set_error("context still has " + std::to_string(ctx->live_circuits.load()) + " circuit(s); destroy them first");
        return GKRX_ERR_STATE;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    dev_free_scratch(ctx->scratch);
    dev_free_batch(ctx->batch);
    dev_bn_free_scratch(ctx->bn);
    for (BnLane &ln : ctx->bn_lanes) {
        if (ln.st) {
            (void)hipStreamSynchronize(ln.st);
            (void)hipStreamDestroy(ln.st);
        }
        dev_bn_free_scratch(ln.bn);
        if (ln.tail_host) (void)hipHostFree(ln.tail_host);
        if (ln.small_host) (void)hipHostFree(ln.small_host);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GKRX_OK;
}

// next_pow_of_2 + _compute_nb_vars (include/utils/myutil.hpp:38-48, circuit.hpp:124-149)
static uint32_t nb_vars_of(uint32_t max_id)
{
    if (max_id == 0) return 0;
    uint32_t p = ((max_id & (max_id - 1)) == 0) ? max_id : (1u << (32 - __builtin_clz(max_id)));
    return (uint32_t)__builtin_ctz(p);
}

// M31(uint32) folds once (include/field/M31.hpp:58-62); the library keeps canonical
// coefficients.  0xffffffff folds to 2^31, where the reference's signed AVX lane
// multiply is not a field multiply, so it is rejected.
static bool canon_coef(uint32_t raw, uint32_t &out)
{
    uint32_t f = (raw & 0x7fffffffu) + (raw >> 31);
    if (f > 0x7fffffffu) return false;
    out = f == 0x7fffffffu ? 0 : f;
    return true;
}

// Shape (_compute_nb_vars, circuit.hpp:124-149) and validation of the gate lists; M31
// coefficients are canonicalized.  No device work.
static int validate_circuit(bool bn254, gkrx_circuit *c)
{
    const uint32_t D = c->D;
    c->n.assign(D, 0);
    c->m.assign(D, 0);
    for (uint32_t l = 0; l < D; l++) {
        uint32_t mo = 0, mi = 0;
        const auto &mu = c->mul[l];
        for (size_t g = 0; g < mu.size() / 4; g++) {
            mi = std::max({mi, mu[4 * g], mu[4 * g + 1]});
            mo = std::max(mo, mu[4 * g + 2]);
        }
        const auto &ad = c->add[l];
        for (size_t g = 0; g < ad.size() / 3; g++) {
            mi = std::max(mi, ad[3 * g]);
            mo = std::max(mo, ad[3 * g + 1]);
        }
        c->n[l] = nb_vars_of(mi);
        c->m[l] = nb_vars_of(mo);
    }
    for (uint32_t l = 0; l < D; l++) {
        const uint32_t n = c->n[l], m = c->m[l];
        if (n < 1 || n > kMaxVars || m < 1 || m > kMaxVars) {
            set_error("layer " + std::to_string(l) + ": need 1..30 input and output variables");
            return GKRX_ERR_CIRCUIT;
        }
        if (l + 1 < D && c->n[l + 1] != m) {
            set_error("layer " + std::to_string(l) + ": output variables differ from the next layer's inputs");
            return GKRX_ERR_CIRCUIT;
        }
        // BN254 keeps the coefficient as read (every u32 is below r)
        auto &mu = c->mul[l];
        for (size_t g = 0; g < mu.size() / 4; g++) {
            if ((mu[4 * g] >> n) || (mu[4 * g + 1] >> n) || (mu[4 * g + 2] >> m) ||
                (!bn254 && !canon_coef(mu[4 * g + 3], mu[4 * g + 3]))) {
                set_error("layer " + std::to_string(l) + ": mul gate index or coefficient out of range");
                return GKRX_ERR_CIRCUIT;
            }
        }
        auto &ad = c->add[l];
        for (size_t g = 0; g < ad.size() / 3; g++) {
            if ((ad[3 * g] >> n) || (ad[3 * g + 1] >> m) || (!bn254 && !canon_coef(ad[3 * g + 2], ad[3 * g + 2]))) {
                set_error("layer " + std::to_string(l) + ": add gate index or coefficient out of range");
                return GKRX_ERR_CIRCUIT;
            }
        }
    }
    return GKRX_OK;
}

static int finish_circuit(gkrx_ctx *ctx, gkrx_circuit *c)
{
    int rc = validate_circuit(ctx->bn254, c);
    if (rc) return rc;
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    rc = dev_upload_circuit(c->dev, ctx->W, c->n, c->m, c->mul, c->add);
    if (rc) return rc;
    c->ctx = ctx;
    ctx->live_circuits++;
    return GKRX_OK;
}

extern "C" int gkrx_circuit_from_arrays(gkrx_ctx *ctx, uint32_t nb_layers, const uint32_t *n_mul,
                                        const uint32_t *const *mul, const uint32_t *n_add,
                                        const uint32_t *const *add, gkrx_circuit **out)
try {
    if (!ctx || !out || nb_layers == 0 || !n_mul || !mul || !n_add || !add) return GKRX_ERR_ARG;
    gkrx_circuit *c = new (std::nothrow) gkrx_circuit;
    if (!c) return GKRX_ERR_NOMEM;
    c->D = nb_layers;
    c->mul.resize(nb_layers);
    c->add.resize(nb_layers);
    for (uint32_t l = 0; l < nb_layers; l++) {
        c->mul[l].assign(mul[l], mul[l] + (size_t)4 * n_mul[l]);
        c->add[l].assign(add[l], add[l] + (size_t)3 * n_add[l]);
    }
    int rc = finish_circuit(ctx, c);
    if (rc) {
        dev_free_circuit(c->dev);
        delete c;
        return rc;
    }
    *out = c;
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

// Bytes from the current position of f to its end (0 when the size is unknown).
static uint64_t bytes_left(FILE *f)
{
    const long here = ftell(f);
    if (here < 0 || fseek(f, 0, SEEK_END) != 0) return 0;
    const long end = ftell(f);
    (void)fseek(f, here, SEEK_SET);
    return end > here ? (uint64_t)(end - here) : 0;
}

// One layer line of the text format: `count` then count x `words` numbers.  The gate list
// grows as gates parse, and the count is checked against what the rest of the file can
// hold (a number and its separator take at least two bytes), so a corrupt count can
// neither wrap a size computation nor reserve memory the file cannot fill.
static const char *read_gate_line(FILE *f, uint32_t words, std::vector<uint32_t> &out)
{
    unsigned long long ng;
    if (fscanf(f, "%llu", &ng) != 1) return "malformed gate count";
    if (ng > 0xffffffffull) return "gate count above 2^32";
    const uint64_t room = bytes_left(f);
    if (ng > room / (2ull * words - 1) + 1) return "gate count larger than the file can hold";
    out.clear();
    out.reserve((size_t)ng * words);
    for (unsigned long long g = 0; g < ng; g++) {
        unsigned long v[4];
        unsigned coef;
        if (words == 4) {
            if (fscanf(f, "%lu%lu%lu%u", &v[0], &v[1], &v[2], &coef) != 4) return "malformed mul gate";
            if ((v[0] | v[1] | v[2]) >> 32) return "gate index above 2^32";
        } else {
            if (fscanf(f, "%lu%lu%u", &v[0], &v[1], &coef) != 3) return "malformed add gate";
            if ((v[0] | v[1]) >> 32) return "gate index above 2^32";
        }
        for (uint32_t w = 0; w + 1 < words; w++) out.push_back((uint32_t)v[w]);
        out.push_back(coef);
    }
    return nullptr;
}

// Parses the two files of the text format into c (D, mul, add); no validation of the
// indices, no device work.
static int parse_text(const char *mul_path, const char *add_path, gkrx_circuit &c)
{
    FILE *f = fopen(mul_path, "r");
    if (!f) {
        set_error(std::string("cannot open ") + mul_path);
        return GKRX_ERR_IO;
    }
    uint32_t nb = 0;
    int ch, last = '\n';
    while ((ch = fgetc(f)) != EOF) {
        if (ch == '\n') nb++;
        last = ch;
    }
    if (last != '\n') nb++;
    rewind(f);
    if (nb == 0) {
        fclose(f);
        set_error("empty mul file");
        return GKRX_ERR_CIRCUIT;
    }
    c.D = nb;
    c.mul.assign(nb, {});
    c.add.assign(nb, {});
    for (uint32_t i = 0; i < nb; i++) {
        // line i -> layers[D - 1 - i]
        if (const char *msg = read_gate_line(f, 4, c.mul[nb - 1 - i])) {
            fclose(f);
            set_error(std::string("mul file: ") + msg);
            return GKRX_ERR_CIRCUIT;
        }
    }
    fclose(f);
    f = fopen(add_path, "r");
    if (!f) {
        set_error(std::string("cannot open ") + add_path);
        return GKRX_ERR_IO;
    }
    for (uint32_t i = 0; i < nb; i++) {
        if (const char *msg = read_gate_line(f, 3, c.add[nb - 1 - i])) {
            fclose(f);
            set_error(std::string("add file: ") + msg);
            return GKRX_ERR_CIRCUIT;
        }
    }
    fclose(f);
    return GKRX_OK;
}

static int load_text_impl(gkrx_ctx *ctx, const char *mul_path, const char *add_path, gkrx_circuit **out)
{
    std::unique_ptr<gkrx_circuit> c(new gkrx_circuit);
    int rc = parse_text(mul_path, add_path, *c);
    if (rc) return rc;
    rc = finish_circuit(ctx, c.get());
    if (rc) {
        dev_free_circuit(c->dev);
        return rc;
    }
    *out = c.release();
    return GKRX_OK;
}

// Circuit::load_extracted_gates (include/circuit/circuit.hpp:152-215)
extern "C" int gkrx_circuit_load_text(gkrx_ctx *ctx, const char *mul_path, const char *add_path, gkrx_circuit **out)
{
    if (!ctx || !mul_path || !add_path || !out) return GKRX_ERR_ARG;
    try {
        return load_text_impl(ctx, mul_path, add_path, out);
    } catch (const std::bad_alloc &) {
        set_error("out of host memory while loading the circuit");
        return GKRX_ERR_NOMEM;
    }
}

extern "C" int gkrx_circuit_parse_text(const char *mul_path, const char *add_path, int field_type,
                                       uint32_t *nb_layers, uint32_t *nb_in_vars, uint32_t *nb_out_vars)
{
    if (!mul_path || !add_path || !nb_layers) return GKRX_ERR_ARG;
    if (field_type != GKRX_FIELD_M31 && field_type != GKRX_FIELD_BN254) return GKRX_ERR_CONFIG;
    try {
        gkrx_circuit c;
        int rc = parse_text(mul_path, add_path, c);
        if (!rc) rc = validate_circuit(field_type == GKRX_FIELD_BN254, &c);
        if (rc) return rc;
        const uint32_t cap = *nb_layers;
        *nb_layers = c.D;
        if (c.D > cap && (nb_in_vars || nb_out_vars)) {
            set_error("shape arrays hold fewer entries than the circuit has layers");
            return GKRX_ERR_ARG;
        }
        for (uint32_t l = 0; l < c.D; l++) {
            if (nb_in_vars) nb_in_vars[l] = c.n[l];
            if (nb_out_vars) nb_out_vars[l] = c.m[l];
        }
        return GKRX_OK;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory while reading the circuit");
        return GKRX_ERR_NOMEM;
    }
}

extern "C" int gkrx_circuit_load_circuit2(gkrx_ctx *ctx, const char *path, int coef_mode, gkrx_circuit **out)
try {
    if (!ctx || !path || !out) return GKRX_ERR_ARG;
    auto *c = new (std::nothrow) gkrx_circuit;
    if (!c) return GKRX_ERR_NOMEM;
    int rc = circuit2_flatten(path, ctx->cfg.field_type, coef_mode, c->mul, c->add);
    if (!rc) {
        c->D = (uint32_t)c->mul.size();
        rc = finish_circuit(ctx, c);
    }
    if (rc) {
        dev_free_circuit(c->dev);
        delete c;
        return rc;
    }
    *out = c;
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_circuit_shape(const gkrx_circuit *c, uint32_t *nb_layers, uint32_t *nb_in_vars,
                                  uint32_t *nb_out_vars)
{
    if (!c) return GKRX_ERR_ARG;
    if (nb_layers) *nb_layers = c->D;
    for (uint32_t l = 0; l < c->D; l++) {
        if (nb_in_vars) nb_in_vars[l] = c->n[l];
        if (nb_out_vars) nb_out_vars[l] = c->m[l];
    }
    return GKRX_OK;
}

extern "C" int gkrx_circuit_write_text(const gkrx_circuit *c, const char *mul_path, const char *add_path)
try {
    if (!c || !mul_path || !add_path) return GKRX_ERR_ARG;
    for (int which = 0; which < 2; which++) {
        FILE *f = fopen(which == 0 ? mul_path : add_path, "w");
        if (!f) {
            set_error("cannot write circuit file");
            return GKRX_ERR_IO;
        }
        for (uint32_t i = 0; i < c->D; i++) {
            const uint32_t l = c->D - 1 - i;
            const auto &g = which == 0 ? c->mul[l] : c->add[l];
            const size_t w = which == 0 ? 4 : 3;
            fprintf(f, "%zu", g.size() / w);
            for (size_t k = 0; k < g.size(); k += w) {
                if (w == 4) fprintf(f, " %u %u %u %u", g[k], g[k + 1], g[k + 2], g[k + 3]);
                else fprintf(f, " %u %u %u", g[k], g[k + 1], g[k + 2]);
            }
            fputc('\n', f);
        }
        fclose(f);
    }
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_circuit_set_witness(gkrx_circuit *c, const uint32_t *lanes, size_t n_u32)
try {
    if (!c || !lanes) return GKRX_ERR_ARG;
    if (c->in_server) {
        set_error("circuit has a server request in flight");
        return GKRX_ERR_STATE;
    }
    const size_t need = ((size_t)1 << c->n[0]) * c->dev.L;
    if (n_u32 != need) {
        set_error("witness must hold 2^nb_input_vars(0) field elements");
        return GKRX_ERR_ARG;
    }
    const bool bn = c->ctx->bn254;
    if (bn) {
        const uint32_t r[8] = GKRX_BN_R;
        for (size_t i = 0; i < n_u32; i += 8)
            if (fr_geq(lanes + i, r)) {
                set_error("witness values must be canonical BN254 scalars (< r)");
                return GKRX_ERR_ARG;
            }
    } else {
        for (size_t i = 0; i < n_u32; i++)
            if (lanes[i] >= 0x7fffffffu) {
                set_error("witness values must be canonical M31 (< 2^31 - 1)");
                return GKRX_ERR_ARG;
            }
    }
    std::lock_guard<std::mutex> lk(c->ctx->mu);
    GKRX_HIP_CHECK(hipSetDevice(c->ctx->device));
    c->witness.assign(lanes, lanes + n_u32);
    GKRX_HIP_CHECK(hipMemcpyAsync(c->dev.values, c->witness.data(), n_u32 * 4, hipMemcpyHostToDevice, c->ctx->stream));
    if (bn) {
        int rc = dev_bn_to_mont(c->dev.values, n_u32 / 8, c->ctx->stream);
        if (rc) return rc;
    }
    GKRX_HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
    c->has_witness = true;
    c->evaluated = false;
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

static int evaluate_locked(gkrx_circuit *c)
{
    if (!c->has_witness) {
        set_error("no witness set");
        return GKRX_ERR_STATE;
    }
    gkrx_ctx *ctx = c->ctx;
    Profiler *pf = ctx->prof.on ? &ctx->prof : nullptr;
    int rc = ctx->bn254 ? dev_bn_forward(c->dev, ctx->stream, pf) : dev_forward(c->dev, ctx->stream, pf);
    if (rc) return rc;
    c->evaluated = true;
    return GKRX_OK;
}

extern "C" int gkrx_circuit_evaluate(gkrx_circuit *c)
{
    if (!c) return GKRX_ERR_ARG;
    if (c->in_server) {
        set_error("circuit has a server request in flight");
        return GKRX_ERR_STATE;
    }
    std::lock_guard<std::mutex> lk(c->ctx->mu);
    GKRX_HIP_CHECK(hipSetDevice(c->ctx->device));
    return evaluate_locked(c);
}

extern "C" int gkrx_circuit_get_layer_values(gkrx_circuit *c, uint32_t layer, uint32_t *out, size_t n_u32)
{
    if (!c || !out || layer > c->D) return GKRX_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->ctx->mu);
    if (!c->evaluated) {
        set_error("circuit not evaluated");
        return GKRX_ERR_STATE;
    }
    const uint32_t vars = layer < c->D ? c->n[layer] : c->m[c->D - 1];
    const size_t need = ((size_t)1 << vars) * c->dev.L;
    if (n_u32 < need) return GKRX_ERR_ARG;
    const uint64_t off = layer < c->D ? c->dev.layers[layer].v_off : c->dev.layers[c->D - 1].out_off;
    GKRX_HIP_CHECK(hipSetDevice(c->ctx->device));
    GKRX_HIP_CHECK(hipMemcpyAsync(out, c->dev.values + off, need * 4, hipMemcpyDeviceToHost, c->ctx->stream));
    GKRX_HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
    if (c->ctx->bn254) {
        // the device keeps Montgomery form; callers get canonical values
        for (size_t i = 0; i < need; i += 8) {
            Fr v;
            memcpy(v.w, out + i, 32);
            v = fr_from_mont(v);
            memcpy(out + i, v.w, 32);
        }
    }
    return GKRX_OK;
}

extern "C" int gkrx_circuit_get_output(gkrx_circuit *c, uint32_t *out, size_t n_u32)
{
    if (!c) return GKRX_ERR_ARG;
    return gkrx_circuit_get_layer_values(c, c->D, out, n_u32);
}

extern "C" int gkrx_circuit_destroy(gkrx_circuit *c)
{
    if (!c) return GKRX_ERR_ARG;
    if (c->in_server) {
        set_error("circuit has a server request in flight");
        return GKRX_ERR_STATE;
    }
    if (c->ctx) {
        std::lock_guard<std::mutex> lk(c->ctx->mu);
        (void)hipSetDevice(c->ctx->device);
        (void)hipStreamSynchronize(c->ctx->stream);
        dev_free_circuit(c->dev);
        c->ctx->live_circuits--;
    }
    delete c;
    return GKRX_OK;
}

extern "C" size_t gkrx_proof_size(const gkrx_ctx *ctx, const gkrx_circuit *c)
{
    if (!ctx || !c) return 0;
    return ((size_t)1 << c->n[0]) * ctx->W * 4 + tail_bytes(ctx, c);
}

// gkrx_prove over BN254 (ctx->mu held, device set)
static int prove_bn254_locked(gkrx_ctx *ctx, gkrx_circuit *c, uint8_t *proof, uint32_t *claimed_v)
{
    const auto t0 = std::chrono::steady_clock::now();
    const size_t tb = tail_bytes(ctx, c);
    int rc = dev_bn_ensure_scratch(ctx->bn, c->dev, tb);
    if (rc) return rc;
    if (!c->evaluated) {
        rc = evaluate_locked(c);  // asynchronous: overlaps the host transcript prefix below
        if (rc) return rc;
    }
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    const size_t commit_len = ((size_t)1 << c->n[0]) * 32;
    const uint32_t mlast = c->m[c->D - 1];
    std::vector<uint32_t> rz1((size_t)8 * mlast);
    uint32_t tstate[10];
    uint8_t grind[32];
    host_transcript_prefix_bn254(kind, reinterpret_cast<const uint8_t *>(c->witness.data()), commit_len, mlast,
                                 (uint32_t)ctx->cfg.grinding_bits, grind, rz1.data(), tstate);
    const auto t1 = std::chrono::steady_clock::now();
    Profiler *pf = ctx->prof.on ? &ctx->prof : nullptr;
    rc = dev_bn_prove(c->dev, ctx->bn, kind, rz1.data(), tstate, grind, 32, ctx->stream, pf);
    if (rc) return rc;
    uint32_t tfinal[10];
    memcpy(proof, c->witness.data(), commit_len);  // RAW commitment bytes (raw.hpp:19-27)
    GKRX_HIP_CHECK(hipMemcpyAsync(proof + commit_len, ctx->bn.tail, tb, hipMemcpyDeviceToHost, ctx->stream));
    GKRX_HIP_CHECK(hipMemcpyAsync(claimed_v, ctx->bn.claims, 32, hipMemcpyDeviceToHost, ctx->stream));
    GKRX_HIP_CHECK(hipMemcpyAsync(tfinal, ctx->bn.tstate, sizeof tfinal, hipMemcpyDeviceToHost, ctx->stream));
    GKRX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto t2 = std::chrono::steady_clock::now();
    if (ctx->prof.on) ctx->prof.settle();
    if (tfinal[0] != tb) {
        set_error("device transcript length mismatch: " + std::to_string(tfinal[0]) + " != " + std::to_string(tb));
        return GKRX_ERR_DEVICE;
    }
    ctx->last.host_prefix_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    ctx->last.device_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    ctx->last.total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    return GKRX_OK;
}

extern "C" int gkrx_prove(gkrx_ctx *ctx, gkrx_circuit *c, uint8_t *proof, size_t proof_cap, size_t *proof_len,
                          uint32_t *claimed_v)
try {
    if (!ctx || !c || !proof || !claimed_v || c->ctx != ctx) return GKRX_ERR_ARG;
    if (c->in_server) {
        set_error("circuit has a server request in flight");
        return GKRX_ERR_STATE;
    }
    const size_t total = gkrx_proof_size(ctx, c);
    if (proof_cap < total) {
        set_error("proof buffer too small");
        return GKRX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->bn254) {
        GKRX_HIP_CHECK(hipSetDevice(ctx->device));
        int rc = prove_bn254_locked(ctx, c, proof, claimed_v);
        if (rc == GKRX_OK) *proof_len = total;
        return rc;
    }
    const auto t0 = std::chrono::steady_clock::now();
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t tb = tail_bytes(ctx, c);
    int rc = dev_ensure_scratch(ctx->scratch, c->dev, ctx->R, tb);
    if (rc) return rc;
    if (!c->evaluated) {
        rc = evaluate_locked(c);  // asynchronous: overlaps the host transcript prefix below
        if (rc) return rc;
    }
    const size_t S = (size_t)ctx->L * 4;
    const size_t commit_len = ((size_t)1 << c->n[0]) * S;
    const uint32_t mlast = c->m[c->D - 1];
    std::vector<uint32_t> rz1((size_t)ctx->R * mlast);
    uint32_t tstate[10];
    uint8_t grind[32];
    host_transcript_prefix(ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0,
                           reinterpret_cast<const uint8_t *>(c->witness.data()), commit_len, ctx->R, mlast,
                           (uint32_t)ctx->cfg.grinding_bits, grind, rz1.data(), tstate);
    const auto t1 = std::chrono::steady_clock::now();
    rc = dev_prove(c->dev, ctx->scratch, ctx->R, ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0, rz1.data(), tstate,
                   grind, 32, ctx->stream, ctx->prof.on ? &ctx->prof : nullptr, ctx->graphs);
    if (rc) return rc;
    uint32_t tfinal[10];
    memcpy(proof, c->witness.data(), commit_len);  // RAW commitment bytes (raw.hpp:19-27)
    GKRX_HIP_CHECK(hipMemcpyAsync(proof + commit_len, ctx->scratch.tail, tb, hipMemcpyDeviceToHost, ctx->stream));
    GKRX_HIP_CHECK(hipMemcpyAsync(claimed_v, ctx->scratch.claims, (size_t)ctx->R * ctx->L * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
    GKRX_HIP_CHECK(hipMemcpyAsync(tfinal, ctx->scratch.tstate, sizeof tfinal, hipMemcpyDeviceToHost, ctx->stream));
    GKRX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto t2 = std::chrono::steady_clock::now();
    if (ctx->prof.on) ctx->prof.settle();
    if (tfinal[0] != tb) {
        set_error("device transcript length mismatch: " + std::to_string(tfinal[0]) + " != " + std::to_string(tb));
        return GKRX_ERR_DEVICE;
    }
    *proof_len = total;
    ctx->last.host_prefix_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    ctx->last.device_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    ctx->last.total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_prove_noalloc(gkrx_ctx *ctx, gkrx_circuit *c, uint8_t *proof, size_t proof_cap,
                                  size_t *proof_len, uint32_t *claimed_v)
{
    if (!ctx || !c || !proof || !proof_len || !claimed_v || c->ctx != ctx) return GKRX_ERR_ARG;
    if (ctx->bn254) {
        set_error("gkrx_prove_noalloc proves M31 circuits");
        return GKRX_ERR_CONFIG;
    }
    if (c->in_server) {
        set_error("circuit has a server request in flight");
        return GKRX_ERR_STATE;
    }
    const size_t total = gkrx_proof_size(ctx, c);
    if (proof_cap < total) {
        set_error("proof buffer too small");
        return GKRX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t tb = tail_bytes(ctx, c);
    if (dev_noalloc_bytes(c->dev, ctx->R, tb) > kNoAllocScratchBytes) {
        set_error("circuit needs " + std::to_string(dev_noalloc_bytes(c->dev, ctx->R, tb)) +
                  " bytes of prover tables, above the " + std::to_string(kNoAllocScratchBytes) +
                  " bytes of gkrx_prove_noalloc (prove it with gkrx_prove)");
        return GKRX_ERR_ARG;
    }
    if (!c->evaluated) {
        int rc = evaluate_locked(c);  // asynchronous: overlaps the host transcript prefix below
        if (rc) return rc;
    }
    const size_t commit_len = ((size_t)1 << c->n[0]) * ctx->L * 4;
    const uint32_t mlast = c->m[c->D - 1];
    uint32_t rz1[3 * kMaxVars], tstate[10], tfinal[10];
    uint8_t grind[32];
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    host_transcript_prefix(kind, reinterpret_cast<const uint8_t *>(c->witness.data()), commit_len, ctx->R, mlast,
                           (uint32_t)ctx->cfg.grinding_bits, grind, rz1, tstate);
    DevScratch view;
    int rc = dev_prove_noalloc(c->dev, ctx->R, kind, rz1, tstate, grind, tb, ctx->device, ctx->stream, view);
    if (rc) return rc;
    memcpy(proof, c->witness.data(), commit_len);  // RAW commitment bytes (raw.hpp:19-27)
    GKRX_HIP_CHECK(hipMemcpyAsync(proof + commit_len, view.tail, tb, hipMemcpyDeviceToHost, ctx->stream));
    GKRX_HIP_CHECK(hipMemcpyAsync(claimed_v, view.claims, (size_t)ctx->R * ctx->L * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
    GKRX_HIP_CHECK(hipMemcpyAsync(tfinal, view.tstate, sizeof tfinal, hipMemcpyDeviceToHost, ctx->stream));
    GKRX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (tfinal[0] != tb) {
        set_error("device transcript length mismatch: " + std::to_string(tfinal[0]) + " != " + std::to_string(tb));
        return GKRX_ERR_DEVICE;
    }
    *proof_len = total;
    return GKRX_OK;
}

namespace {

struct BatchPrefix {
    std::vector<uint32_t> rz1;
    uint32_t tstate[10];
    uint8_t grind[32];
};

// The transcript prefixes of a batch (independent serial hash chains) on host threads, in
// the order in which the device needs them.  The first prefixes needed run one per thread
// (a scalar sponge finishes one commitment soonest); with the eight-lane Keccak the rest
// run as groups of up to eight commitments of one length, each led by the earliest-needed
// item not yet taken and filled with the next-needed items of its length (eight sponges
// take about 2.7 times as long as one on a core).
class PrefixRunner final : public BatchPrefixes {
  public:
    PrefixRunner(gkrx_ctx *ctx, gkrx_circuit *const *circuits, uint32_t n, int kind, std::vector<BatchPrefix> &pre)
        : ctx_(ctx), circuits_(circuits), n_(n), kind_(kind), pre_(pre), done_(n, 0)
    {
    }
    ~PrefixRunner() override { join(); }

    void start(const std::vector<uint32_t> &order) override
    {
        t0_ = std::chrono::steady_clock::now();
        const bool x8 = kind_ == 1 && host_keccak_x8();
        const unsigned nt = std::max(1u, std::min(16u, usable_cpus()));
        std::vector<uint8_t> taken(n_, 0);
        for (size_t k = 0; k < order.size(); k++) {
            const uint32_t lead = order[k];
            if (taken[lead]) continue;
            std::vector<uint32_t> g{lead};
            taken[lead] = 1;
            for (size_t q = k + 1; x8 && k >= nt && q < order.size() && g.size() < 8; q++)
                if (!taken[order[q]] && circuits_[order[q]]->n[0] == circuits_[lead]->n[0]) {
                    g.push_back(order[q]);
                    taken[order[q]] = 1;
                }
            groups_.push_back(std::move(g));
        }
        for (size_t t = 0; t < std::min<size_t>(nt, groups_.size()); t++)
            pool_.emplace_back([this] {
                for (size_t g; (g = next_++) < groups_.size();) {
                    run(groups_[g]);
                    {
                        std::lock_guard<std::mutex> lk(mu_);
                        for (uint32_t i : groups_[g]) done_[i] = 1;
                        ndone_ += groups_[g].size();
                        if (ndone_ == n_) t1_ = std::chrono::steady_clock::now();
                    }
                    cv_.notify_all();
                }
            });
    }

    void wait(uint32_t i) override
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return done_[i] != 0; });
    }

    // joins the threads; returns their wall time in ms (0 if they never started)
    double join()
    {
        for (auto &th : pool_) th.join();
        pool_.clear();
        if (ndone_ != n_ || n_ == 0) return 0;
        return std::chrono::duration<double, std::milli>(t1_ - t0_).count();
    }

  private:
    void run(const std::vector<uint32_t> &g)
    {
        const size_t S = (size_t)ctx_->L * 4;
        const uint32_t gb = (uint32_t)ctx_->cfg.grinding_bits;
        gkrx_circuit *c0 = circuits_[g[0]];
        const size_t len = ((size_t)1 << c0->n[0]) * S;
        if (g.size() == 1) {
            BatchPrefix &p = pre_[g[0]];
            host_transcript_prefix(kind_, reinterpret_cast<const uint8_t *>(c0->witness.data()), len, ctx_->R,
                                   c0->m[c0->D - 1], gb, p.grind, p.rz1.data(), p.tstate);
            return;
        }
        const uint8_t *commits[8];
        uint32_t ml[8], *rz[8];
        uint8_t grind[8][32];
        uint32_t ts[8][10];
        for (size_t l = 0; l < g.size(); l++) {
            gkrx_circuit *c = circuits_[g[l]];
            commits[l] = reinterpret_cast<const uint8_t *>(c->witness.data());
            ml[l] = c->m[c->D - 1];
            rz[l] = pre_[g[l]].rz1.data();
        }
        host_transcript_prefix_x8((uint32_t)g.size(), commits, len, ctx_->R, ml, gb, grind, rz, ts);
        for (size_t l = 0; l < g.size(); l++) {
            memcpy(pre_[g[l]].grind, grind[l], 32);
            memcpy(pre_[g[l]].tstate, ts[l], 40);
        }
    }

    gkrx_ctx *ctx_;
    gkrx_circuit *const *circuits_;
    uint32_t n_;
    int kind_;
    std::vector<BatchPrefix> &pre_;
    std::vector<uint8_t> done_;
    size_t ndone_ = 0;
    std::vector<std::vector<uint32_t>> groups_;
    std::atomic<size_t> next_{0};
    std::vector<std::thread> pool_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::chrono::steady_clock::time_point t0_, t1_;
};

}  // namespace

// BN254 batches: the host prefixes on host threads, then the proofs on up to kBnBatchLanes
// streams at once (each proof's rounds are chains of short launches and hashes, so proofs
// side by side fill the GPU that one leaves idle).  Proof i is gkrx_prove's bytes for
// circuits[i].  ctx->mu is not held on entry.
static int prove_batch_bn254(gkrx_ctx *ctx, uint32_t n, gkrx_circuit *const *circuits, uint8_t *const *proofs,
                             size_t *proof_lens, uint32_t *const *claimed_v)
{
    std::lock_guard<std::mutex> lk(ctx->mu);
    const auto t0 = std::chrono::steady_clock::now();
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    struct Pre {
        std::vector<uint32_t> rz1;
        uint32_t tstate[10];
        uint8_t grind[32];
    };
    std::vector<Pre> pre(n);
    parallel_for(n, 16, [&](size_t i) {
        gkrx_circuit *c = circuits[i];
        const uint32_t mlast = c->m[c->D - 1];
        pre[i].rz1.resize((size_t)8 * mlast);
        host_transcript_prefix_bn254(kind, reinterpret_cast<const uint8_t *>(c->witness.data()),
                                     commit_bytes(ctx, c), mlast, (uint32_t)ctx->cfg.grinding_bits, pre[i].grind,
                                     pre[i].rz1.data(), pre[i].tstate);
    });
    const auto t1 = std::chrono::steady_clock::now();
    const uint32_t nl = std::min<uint32_t>(n, kBnBatchLanes);
    if (ctx->bn_lanes.size() < nl) ctx->bn_lanes.resize(nl);
    for (uint32_t q = 0; q < nl; q++) {
        BnLane &ln = ctx->bn_lanes[q];
        if (ln.st) GKRX_HIP_CHECK(hipStreamSynchronize(ln.st));  // a failed batch may have left work
        ln.item = -1;
        if (!ln.st) GKRX_HIP_CHECK(hipStreamCreateWithFlags(&ln.st, hipStreamNonBlocking));
        if (!ln.small_host) GKRX_HIP_CHECK(hipHostMalloc(&ln.small_host, 32 * 4, hipHostMallocDefault));
    }
    // waits for the lane's proof and hands its tail and claims out
    auto finish = [&](BnLane &ln) -> int {
        if (ln.item < 0) return GKRX_OK;
        const uint32_t i = (uint32_t)ln.item;
        ln.item = -1;
        GKRX_HIP_CHECK(hipStreamSynchronize(ln.st));
        gkrx_circuit *c = circuits[i];
        const size_t tb = tail_bytes(ctx, c);
        if (ln.small_host[8] != tb) {
            set_error("batch item " + std::to_string(i) + ": device transcript length " +
                      std::to_string(ln.small_host[8]) + " != " + std::to_string(tb));
            return GKRX_ERR_DEVICE;
        }
        memcpy(proofs[i], c->witness.data(), commit_bytes(ctx, c));  // RAW commitment (raw.hpp:19-27)
        memcpy(proofs[i] + commit_bytes(ctx, c), ln.tail_host, tb);
        memcpy(claimed_v[i], ln.small_host, 32);
        proof_lens[i] = commit_bytes(ctx, c) + tb;
        c->evaluated = true;
        return GKRX_OK;
    };
    int rc = GKRX_OK;
    for (uint32_t i = 0; i < n && rc == GKRX_OK; i++) {
        BnLane &ln = ctx->bn_lanes[i % nl];
        if ((rc = finish(ln))) break;
        gkrx_circuit *c = circuits[i];
        const size_t tb = tail_bytes(ctx, c);
        if (tb > ln.tail_cap) {
            if (ln.tail_host) (void)hipHostFree(ln.tail_host);
            ln.tail_host = nullptr;
            ln.tail_cap = 0;
            GKRX_HIP_CHECK(hipHostMalloc(&ln.tail_host, tb, hipHostMallocDefault));
            ln.tail_cap = tb;
        }
        if ((rc = dev_bn_ensure_scratch(ln.bn, c->dev, tb))) break;
        if ((rc = dev_bn_forward(c->dev, ln.st, nullptr))) break;
        if ((rc = dev_bn_prove(c->dev, ln.bn, kind, pre[i].rz1.data(), pre[i].tstate, pre[i].grind, 32, ln.st,
                               nullptr)))
            break;
        GKRX_HIP_CHECK(hipMemcpyAsync(ln.tail_host, ln.bn.tail, tb, hipMemcpyDeviceToHost, ln.st));
        GKRX_HIP_CHECK(hipMemcpyAsync(ln.small_host, ln.bn.claims, 32, hipMemcpyDeviceToHost, ln.st));
        GKRX_HIP_CHECK(hipMemcpyAsync(ln.small_host + 8, ln.bn.tstate, 40, hipMemcpyDeviceToHost, ln.st));
        ln.item = i;
    }
    for (uint32_t q = 0; q < nl; q++) {
        const int r2 = finish(ctx->bn_lanes[q]);
        if (rc == GKRX_OK) rc = r2;
    }
    if (rc) return rc;
    const auto t2 = std::chrono::steady_clock::now();
    ctx->last.host_prefix_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    ctx->last.device_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    ctx->last.total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    return GKRX_OK;
}

extern "C" int gkrx_prove_batch(gkrx_ctx *ctx, uint32_t n, gkrx_circuit *const *circuits, uint8_t *const *proofs,
                                const size_t *proof_caps, size_t *proof_lens, uint32_t *const *claimed_v)
try {
    if (!ctx || (n && (!circuits || !proofs || !proof_caps || !proof_lens || !claimed_v))) return GKRX_ERR_ARG;
    for (uint32_t i = 0; i < n; i++) {
        gkrx_circuit *c = circuits[i];
        if (!c || c->ctx != ctx || !proofs[i] || !claimed_v[i]) return GKRX_ERR_ARG;
        if (c->in_server) {
            set_error("batch item " + std::to_string(i) + " has a server request in flight");
            return GKRX_ERR_STATE;
        }
        if (!c->has_witness) {
            set_error("batch item " + std::to_string(i) + " has no witness");
            return GKRX_ERR_STATE;
        }
        if (proof_caps[i] < gkrx_proof_size(ctx, c)) {
            set_error("batch item " + std::to_string(i) + ": proof buffer too small");
            return GKRX_ERR_ARG;
        }
        if (!ctx->bn254 && !dev_batch_supported(c->dev)) {
            set_error("batch item " + std::to_string(i) + ": layers wider than 2^20 (prove it with gkrx_prove)");
            return GKRX_ERR_CIRCUIT;
        }
    }
    if (n == 0) return GKRX_OK;
    if (ctx->bn254) return prove_batch_bn254(ctx, n, circuits, proofs, proof_lens, claimed_v);
    std::lock_guard<std::mutex> lk(ctx->mu);
    const auto t0 = std::chrono::steady_clock::now();
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    const size_t S = (size_t)ctx->L * 4;
    std::vector<BatchPrefix> pre(n);
    for (uint32_t i = 0; i < n; i++) pre[i].rz1.resize((size_t)ctx->R * circuits[i]->m[circuits[i]->D - 1]);
    std::vector<BatchItem> items(n);
    for (uint32_t i = 0; i < n; i++) {
        gkrx_circuit *c = circuits[i];
        const size_t commit_len = ((size_t)1 << c->n[0]) * S;
        items[i] = BatchItem{&c->dev, pre[i].rz1.data(), pre[i].tstate, pre[i].grind, tail_bytes(ctx, c),
                             proofs[i] + commit_len};
    }
    // the prefixes run on host threads while the batch is enqueued and proved
    PrefixRunner runner(ctx, circuits, n, kind, pre);
    const auto t1 = std::chrono::steady_clock::now();
    int rc = dev_prove_batch(items, ctx->R, kind, ctx->batch, ctx->stream, ctx->prof.on ? &ctx->prof : nullptr,
                             &runner);
    const double prefix_ms = runner.join();
    if (rc) return rc;
    // while the device proves: the commitments (the input layers) into the proof buffers
    parallel_for(n, 16, [&](size_t i) {
        memcpy(proofs[i], circuits[i]->witness.data(), ((size_t)1 << circuits[i]->n[0]) * S);
    });
    GKRX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    rc = dev_batch_collect(items, ctx->batch);
    if (rc) return rc;
    const uint32_t *cl = ctx->batch.claims_host;
    if (!cl) return GKRX_ERR_NOMEM;
    for (uint32_t i = 0; i < n; i++) {
        memcpy(claimed_v[i], cl + (size_t)i * ctx->R * ctx->L, (size_t)ctx->R * ctx->L * 4);
        proof_lens[i] = gkrx_proof_size(ctx, circuits[i]);
        circuits[i]->evaluated = true;
    }
    if (ctx->prof.on) ctx->prof.settle();
    const auto t2 = std::chrono::steady_clock::now();
    ctx->last.host_prefix_ms = prefix_ms;
    ctx->last.device_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    ctx->last.total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_ctx_set_batch_engine(gkrx_ctx *ctx, int engine, uint32_t block)
{
    if (!ctx) return GKRX_ERR_ARG;
    if ((engine != GKRX_BATCH_PER_PROOF && engine != GKRX_BATCH_LOCKSTEP && engine != GKRX_BATCH_LOCKSTEP_GROUPS) ||
        (block != 64 && block != 128 && block != 256 && block != 512)) {
        set_error("batch engine must be GKRX_BATCH_PER_PROOF, GKRX_BATCH_LOCKSTEP or GKRX_BATCH_LOCKSTEP_GROUPS with "
                  "block 64, 128, 256 or 512");
        return GKRX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->batch.engine = engine;
    ctx->batch.block = block;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_get_batch_engine(const gkrx_ctx *ctx, int *engine, uint32_t *block)
{
    if (!ctx) return GKRX_ERR_ARG;
    if (engine) *engine = ctx->batch.engine;
    if (block) *block = ctx->batch.block;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_set_graphs(gkrx_ctx *ctx, int on)
{
    if (!ctx || (on != 0 && on != 1)) {
        set_error("graphs: 0 or 1");
        return GKRX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->graphs = on != 0;
    if (!ctx->graphs) {
        GKRX_HIP_CHECK(hipSetDevice(ctx->device));
        GKRX_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // a replay may still run
        dev_drop_graph(ctx->scratch);
    }
    return GKRX_OK;
}

extern "C" int gkrx_ctx_get_graphs(const gkrx_ctx *ctx, int *on)
{
    if (!ctx || !on) return GKRX_ERR_ARG;
    *on = ctx->graphs ? 1 : 0;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_set_batch_lanes(gkrx_ctx *ctx, uint32_t lanes)
{
    if (!ctx) return GKRX_ERR_ARG;
    if (lanes != 1 && lanes != 2 && lanes != 4) {
        set_error("lanes per thread must be 1, 2 or 4");
        return GKRX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->batch.lanes = lanes;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_get_batch_lanes(const gkrx_ctx *ctx, uint32_t *lanes)
{
    if (!ctx || !lanes) return GKRX_ERR_ARG;
    *lanes = ctx->batch.lanes;
    return GKRX_OK;
}

extern "C" int gkrx_batch_autotune(gkrx_ctx *ctx, uint32_t n, gkrx_circuit *const *circuits, int *engine,
                                   uint32_t *block, uint32_t *lanes, double *ms_each)
try {
    if (!ctx || n == 0 || !circuits) return GKRX_ERR_ARG;
    if (ctx->bn254) {
        set_error("the batch engines prove M31 circuits");
        return GKRX_ERR_CONFIG;
    }
    for (uint32_t i = 0; i < n; i++) {
        gkrx_circuit *c = circuits[i];
        if (!c || c->ctx != ctx) return GKRX_ERR_ARG;
        if (c->in_server || !c->has_witness || !dev_batch_supported(c->dev)) {
            set_error("autotune item " + std::to_string(i) + " has no witness, is in a server or is too wide");
            return GKRX_ERR_STATE;
        }
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    // each candidate proves the batch as gkrx_prove_batch does: the host transcript prefixes
    // run on host threads while the batch is enqueued, so a schedule that needs prefixes
    // early pays for the wait here too (the proof bytes go to scratch buffers)
    std::vector<BatchPrefix> pre(n);
    std::vector<std::vector<uint8_t>> tails(n);
    std::vector<BatchItem> items(n);
    for (uint32_t i = 0; i < n; i++) {
        gkrx_circuit *c = circuits[i];
        pre[i].rz1.resize((size_t)ctx->R * c->m[c->D - 1]);
        tails[i].resize(tail_bytes(ctx, c));
        items[i] = BatchItem{&c->dev, pre[i].rz1.data(), pre[i].tstate, pre[i].grind, tails[i].size(), tails[i].data()};
    }
    struct Cand {
        int engine;
        uint32_t block, lanes;
    };
    // the one-workgroup-per-proof engine, then the lockstep engine over block size x lanes
    std::vector<Cand> cands = {{GKRX_BATCH_PER_PROOF, 256, 4}};
    for (int e : {GKRX_BATCH_LOCKSTEP, GKRX_BATCH_LOCKSTEP_GROUPS})
        for (uint32_t b : {64u, 128u, 256u, 512u})
            for (uint32_t v : {1u, 2u, 4u}) cands.push_back({e, b, v});
    const int saved_engine = ctx->batch.engine;
    const uint32_t saved_block = ctx->batch.block, saved_lanes = ctx->batch.lanes;
    std::vector<double> times(cands.size());
    for (size_t k = 0; k < cands.size(); k++) {
        ctx->batch.engine = cands[k].engine;
        ctx->batch.block = cands[k].block;
        ctx->batch.lanes = cands[k].lanes;
        double ms = 1e300;
        for (int rep = 0; rep < 3; rep++) {  // the first run allocates and loads code; the best of the other two
            const auto t0 = std::chrono::steady_clock::now();
            PrefixRunner runner(ctx, circuits, n, kind, pre);
            int rc = dev_prove_batch(items, ctx->R, kind, ctx->batch, ctx->stream, nullptr, &runner);
            runner.join();
            if (rc == GKRX_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = GKRX_ERR_DEVICE;
            if (rc) {
                ctx->batch.engine = saved_engine;
                ctx->batch.block = saved_block;
                ctx->batch.lanes = saved_lanes;
                return rc;
            }
            const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (rep > 0) ms = std::min(ms, t);
        }
        times[k] = ms;
        if (ms_each) ms_each[k] = ms;
    }
    // settings within 1% of the fastest are tied (run-to-run noise on a batch is about that
    // much); the first of them in the fixed order above wins, so the pick repeats
    const double best = *std::min_element(times.begin(), times.end());
    Cand pick{saved_engine, saved_block, saved_lanes};
    for (size_t k = 0; k < cands.size(); k++)
        if (times[k] <= best * 1.01) {
            pick = cands[k];
            break;
        }
    ctx->batch.engine = pick.engine;
    ctx->batch.block = pick.block;
    ctx->batch.lanes = pick.lanes;
    if (engine) *engine = pick.engine;
    if (block) *block = pick.block;
    if (lanes) *lanes = pick.lanes;
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_verify(gkrx_ctx *ctx, const gkrx_circuit *c, const uint32_t *claimed_v, const uint8_t *proof,
                           size_t proof_len)
try {
    if (!ctx || !c || !claimed_v || !proof) return GKRX_ERR_ARG;
    HostCircuitView v{c->D, c->n.data(), c->m.data(), c->mul, c->add};
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    if (ctx->bn254) return host_verify_bn254(v, kind, (uint32_t)ctx->cfg.grinding_bits, claimed_v, proof, proof_len);
    return host_verify(v, ctx->L, ctx->R, kind, (uint32_t)ctx->cfg.grinding_bits, claimed_v, proof, proof_len);
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_verify_arrays(const gkrx_config *cfg, uint32_t nb_layers, const uint32_t *n_mul,
                                  const uint32_t *const *mul, const uint32_t *n_add, const uint32_t *const *add,
                                  const uint32_t *claimed_v, const uint8_t *proof, size_t proof_len)
try {
    if (!cfg || nb_layers == 0 || !n_mul || !mul || !n_add || !add || !claimed_v || !proof) return GKRX_ERR_ARG;
    const bool bn = cfg->field_type == GKRX_FIELD_BN254;
    if ((cfg->field_type != GKRX_FIELD_M31 && !bn) || (cfg->fs_hash != GKRX_FS_SHA256 && cfg->fs_hash != GKRX_FS_KECCAK256) ||
        (!bn && cfg->nb_parallel != 8 && cfg->nb_parallel != 16) || cfg->grinding_bits < 0 || cfg->grinding_bits > 24) {
        set_error("unsupported config");
        return GKRX_ERR_CONFIG;
    }
    gkrx_circuit c;
    c.D = nb_layers;
    c.mul.resize(nb_layers);
    c.add.resize(nb_layers);
    for (uint32_t l = 0; l < nb_layers; l++) {
        c.mul[l].assign(mul[l], mul[l] + (size_t)4 * n_mul[l]);
        c.add[l].assign(add[l], add[l] + (size_t)3 * n_add[l]);
    }
    int rc = validate_circuit(bn, &c);
    if (rc) return rc;
    HostCircuitView v{c.D, c.n.data(), c.m.data(), c.mul, c.add};
    const int kind = cfg->fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    if (bn) return host_verify_bn254(v, kind, (uint32_t)cfg->grinding_bits, claimed_v, proof, proof_len);
    const int R = gkrx_config_num_repetitions(cfg);
    if (R < 1 || R > 3) return GKRX_ERR_CONFIG;
    return host_verify(v, (uint32_t)cfg->nb_parallel, (uint32_t)R, kind, (uint32_t)cfg->grinding_bits, claimed_v,
                       proof, proof_len);
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

// Proof::write_and_compress / read_proof file layout (include/fiat_shamir/transcript.hpp:54-85)
extern "C" int gkrx_proof_write(const char *path, const uint8_t *proof, size_t proof_len)
{
    if (!path || (!proof && proof_len)) return GKRX_ERR_ARG;
    if (proof_len > 0xffffffffull) {
        set_error("a proof file holds at most 2^32 - 1 bytes");
        return GKRX_ERR_ARG;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error(std::string("cannot write ") + path);
        return GKRX_ERR_IO;
    }
    const uint32_t n = (uint32_t)proof_len;
    const bool ok = fwrite(&n, 4, 1, f) == 1 && (proof_len == 0 || fwrite(proof, 1, proof_len, f) == proof_len);
    if (fclose(f) != 0 || !ok) {
        set_error(std::string("write failed: ") + path);
        return GKRX_ERR_IO;
    }
    return GKRX_OK;
}

extern "C" int gkrx_proof_read(const char *path, uint8_t *proof, size_t proof_cap, size_t *proof_len)
{
    if (!path || !proof_len) return GKRX_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error(std::string("cannot open ") + path);
        return GKRX_ERR_IO;
    }
    uint32_t n = 0;
    if (fread(&n, 4, 1, f) != 1) {
        fclose(f);
        set_error("truncated proof file");
        return GKRX_ERR_IO;
    }
    *proof_len = n;
    if (!proof) {
        fclose(f);
        return GKRX_OK;
    }
    if (proof_cap < n) {
        fclose(f);
        set_error("proof buffer too small");
        return GKRX_ERR_ARG;
    }
    const bool ok = fread(proof, 1, n, f) == n;
    fclose(f);
    if (!ok) {
        set_error("truncated proof file");
        return GKRX_ERR_IO;
    }
    return GKRX_OK;
}

extern "C" int gkrx_last_timing(const gkrx_ctx *ctx, double *ms3)
{
    if (!ctx || !ms3) return GKRX_ERR_ARG;
    ms3[0] = ctx->last.host_prefix_ms;
    ms3[1] = ctx->last.device_ms;
    ms3[2] = ctx->last.total_ms;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_profile(gkrx_ctx *ctx, int enable)
{
    if (!ctx) return GKRX_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->prof.reset();
    ctx->prof.on = enable != 0;
    ctx->prof.mask = ~0ull;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_profile_classes(gkrx_ctx *ctx, uint64_t mask)
{
    if (!ctx) return GKRX_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->prof.reset();
    ctx->prof.on = mask != 0;
    ctx->prof.mask = mask;
    return GKRX_OK;
}

extern "C" int gkrx_ctx_kernel_stat(gkrx_ctx *ctx, int idx, const char **name, uint64_t *launches, double *ms)
{
    if (!ctx || idx < 0 || idx >= kKClassCount) return GKRX_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->prof.collect();  // every prove that recorded events has synchronized its streams
    if (name) *name = kernel_class_name(idx);
    if (launches) *launches = ctx->prof.count[idx];
    if (ms) *ms = ctx->prof.ms[idx];
    return GKRX_OK;
}

extern "C" int gkrx_hash(int fs_hash, const uint8_t *in, size_t len, uint8_t out[32])
{
    if ((!in && len) || !out) return GKRX_ERR_ARG;
    if (fs_hash != GKRX_FS_SHA256 && fs_hash != GKRX_FS_KECCAK256) return GKRX_ERR_CONFIG;
    host_hash(fs_hash == GKRX_FS_KECCAK256 ? 1 : 0, in, len, out);
    return GKRX_OK;
}

extern "C" int gkrx_hash_many(int fs_hash, uint32_t n, const uint8_t *const *in, size_t len, uint8_t *out)
try {
    if ((n && (!in || !out))) return GKRX_ERR_ARG;
    for (uint32_t i = 0; i < n; i++)
        if (!in[i] && len) return GKRX_ERR_ARG;
    if (fs_hash == GKRX_FS_KECCAK256) {
        host_keccak256_many(n, in, len, out);
        return host_keccak_x8() ? 1 : 0;
    }
    if (fs_hash != GKRX_FS_SHA256) return GKRX_ERR_CONFIG;
    for (uint32_t i = 0; i < n; i++) host_sha256(in[i], len, out + 32 * (size_t)i);
    return 0;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}
