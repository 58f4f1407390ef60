// Runs the library's host paths -- single proves, the allocation-free prove, every batch
// engine (schedule built, reused and rebuilt), the autotuner, BN254 proves and batches, and
// the server -- against the stubbed HIP runtime (hip_stub.cpp), built with AddressSanitizer
// and UndefinedBehaviorSanitizer.  No kernel runs, so the device transcript never reaches
// the expected length: every prove ends in GKRX_ERR_DEVICE, which is the expected, orderly
// outcome here; any other failure, a sanitizer report or a stub abort fails the run.
#include <gkrx.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

static int fails = 0;
static void expect(const char *what, int rc, int want)
{
    if (rc != want) {
        std::printf("FAIL %s: rc %d, expected %d (%s)\n", what, rc, want, gkrx_last_error());
        fails++;
    } else if (rc == GKRX_ERR_DEVICE) {
        std::printf("%s: %s\n", what, gkrx_last_error());  // where the unrun kernels were noticed
    }
}

struct Circ {
    gkrx_circuit *c = nullptr;
    std::vector<uint32_t> witness;
};

static Circ make(gkrx_ctx *ctx, uint64_t seed, uint32_t D, uint32_t k, bool bn254, uint32_t L)
{
    std::vector<std::vector<uint32_t>> mul(D), add(D);
    std::vector<uint32_t> nm(D), na(D);
    std::vector<const uint32_t *> mp(D), ap(D);
    for (uint32_t l = 0; l < D; l++) {
        mul[l].resize((size_t)4 << k);
        add[l].resize((size_t)3 << k);
        gkrx_synth_layer(seed, l, k, 700, 1, mul[l].data(), &nm[l], add[l].data(), &na[l]);
        mp[l] = mul[l].data();
        ap[l] = add[l].data();
    }
    Circ r;
    expect("from_arrays", gkrx_circuit_from_arrays(ctx, D, nm.data(), mp.data(), na.data(), ap.data(), &r.c), 0);
    if (bn254) {
        r.witness.resize((size_t)8 << k);
        gkrx_synth_witness_bn254(seed, 1ull << k, r.witness.data());
    } else {
        r.witness.resize((size_t)L << k);
        gkrx_synth_witness_m31(seed, 1ull << k, L, r.witness.data());
    }
    expect("set_witness", gkrx_circuit_set_witness(r.c, r.witness.data(), r.witness.size()), 0);
    return r;
}

static void batch(gkrx_ctx *ctx, std::vector<Circ> &cs, const char *what)
{
    const uint32_t n = (uint32_t)cs.size();
    std::vector<gkrx_circuit *> h(n);
    std::vector<std::vector<uint8_t>> proofs(n);
    std::vector<std::vector<uint32_t>> claims(n, std::vector<uint32_t>(3 * 16));
    std::vector<uint8_t *> pp(n);
    std::vector<size_t> caps(n), lens(n);
    std::vector<uint32_t *> cp(n);
    for (uint32_t i = 0; i < n; i++) {
        h[i] = cs[i].c;
        proofs[i].resize(gkrx_proof_size(ctx, cs[i].c));
        pp[i] = proofs[i].data();
        caps[i] = proofs[i].size();
        cp[i] = claims[i].data();
    }
    expect(what, gkrx_prove_batch(ctx, n, h.data(), pp.data(), caps.data(), lens.data(), cp.data()), GKRX_ERR_DEVICE);
}

int main()
{
    // M31 x16, Keccak-256
    gkrx_config cfg;
    gkrx_config_default(&cfg);
    cfg.nb_parallel = 16;
    cfg.fs_hash = GKRX_FS_KECCAK256;
    gkrx_ctx *ctx = nullptr;
    expect("ctx_create", gkrx_ctx_create(&cfg, 0, &ctx), 0);
    std::vector<Circ> cs;
    const uint32_t shapes[][2] = {{3, 6}, {5, 9}, {2, 12}, {8, 4}, {4, 13}, {6, 10}, {3, 14}, {2, 2}, {7, 11}, {4, 8}};
    for (uint32_t i = 0; i < 10; i++) cs.push_back(make(ctx, 100 + i, shapes[i][0], shapes[i][1], false, 16));
    {
        std::vector<uint8_t> proof(gkrx_proof_size(ctx, cs[1].c));
        std::vector<uint32_t> cl(48);
        size_t len = 0;
        expect("prove", gkrx_prove(ctx, cs[1].c, proof.data(), proof.size(), &len, cl.data()), GKRX_ERR_DEVICE);
        expect("prove_noalloc", gkrx_prove_noalloc(ctx, cs[1].c, proof.data(), proof.size(), &len, cl.data()),
               GKRX_ERR_DEVICE);
    }
    const int engines[] = {GKRX_BATCH_PER_PROOF, GKRX_BATCH_LOCKSTEP, GKRX_BATCH_LOCKSTEP_GROUPS};
    for (int e : engines)
        for (uint32_t block : {64u, 256u})
            for (uint32_t lanes : {1u, 4u}) {
                expect("set_engine", gkrx_ctx_set_batch_engine(ctx, e, block), 0);
                expect("set_lanes", gkrx_ctx_set_batch_lanes(ctx, lanes), 0);
                batch(ctx, cs, "batch");  // builds the schedule
                batch(ctx, cs, "batch again");  // reuses it
                std::vector<Circ> sub(cs.begin() + 2, cs.begin() + 7);
                batch(ctx, sub, "batch subset");  // rebuilds it
            }
    expect("profile", gkrx_ctx_profile(ctx, 1), 0);
    batch(ctx, cs, "batch profiled");
    expect("profile off", gkrx_ctx_profile(ctx, 0), 0);
    {
        std::vector<gkrx_circuit *> h;
        for (auto &c : cs) h.push_back(c.c);
        int e = 0;
        uint32_t b = 0, l = 0;
        std::vector<double> ms(25);
        const int rc = gkrx_batch_autotune(ctx, (uint32_t)h.size(), h.data(), &e, &b, &l, ms.data());
        if (rc != GKRX_ERR_DEVICE && rc != 0) {
            std::printf("FAIL autotune rc %d (%s)\n", rc, gkrx_last_error());
            fails++;
        }
    }
    // the server: every request fails in the orderly way, the server drains and goes
    {
        gkrx_server *srv = nullptr;
        expect("server_create", gkrx_server_create(ctx, 1ull << 30, 2, &srv), 0);
        std::vector<std::vector<uint8_t>> proofs(cs.size());
        std::vector<std::vector<uint32_t>> cl(cs.size(), std::vector<uint32_t>(48));
        std::vector<uint64_t> tickets(cs.size());
        for (size_t i = 0; i < cs.size(); i++) {
            proofs[i].resize(gkrx_proof_size(ctx, cs[i].c));
            expect("submit", gkrx_server_submit(srv, cs[i].c, proofs[i].data(), proofs[i].size(), cl[i].data(),
                                                &tickets[i]),
                   0);
        }
        for (size_t i = 0; i < cs.size(); i++) {
            size_t len = 0;
            double split[6];
            const int rc = gkrx_server_wait_split(srv, tickets[i], -1, &len, split);
            if (rc >= 0) {
                std::printf("FAIL server request %zu: rc %d, expected a device error\n", i, rc);
                fails++;
            }
        }
        expect("server_destroy", gkrx_server_destroy(srv), 0);
    }
    for (auto &c : cs) expect("circuit_destroy", gkrx_circuit_destroy(c.c), 0);
    expect("ctx_destroy", gkrx_ctx_destroy(ctx), 0);

    // BN254: single proves and a batch over the context's parallel lanes
    gkrx_config bcfg;
    gkrx_config_default(&bcfg);
    bcfg.field_type = GKRX_FIELD_BN254;
    bcfg.fs_hash = GKRX_FS_KECCAK256;
    expect("bn ctx_create", gkrx_ctx_create(&bcfg, 0, &ctx), 0);
    std::vector<Circ> bs;
    for (uint32_t i = 0; i < 6; i++) bs.push_back(make(ctx, 200 + i, 2 + i % 3, 6 + 2 * i, true, 1));
    {
        std::vector<uint8_t> proof(gkrx_proof_size(ctx, bs[3].c));
        std::vector<uint32_t> cl(8);
        size_t len = 0;
        expect("bn prove", gkrx_prove(ctx, bs[3].c, proof.data(), proof.size(), &len, cl.data()), GKRX_ERR_DEVICE);
    }
    batch(ctx, bs, "bn batch");
    batch(ctx, bs, "bn batch again");
    for (auto &c : bs) expect("bn circuit_destroy", gkrx_circuit_destroy(c.c), 0);
    expect("bn ctx_destroy", gkrx_ctx_destroy(ctx), 0);
    std::printf("%s: %d failure(s)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
