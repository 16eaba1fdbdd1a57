// sanitize_host.cpp -- drives the library's host code under AddressSanitizer +
// UndefinedBehaviorSanitizer or ThreadSanitizer (host code only; device code
// is built normally).  Built with the whole library from source by
// tests/cpp/Makefile and run by tests/test_sanitizers.py.
//
// Sections (each prints one "ok <name>" line):
//   ingest   16-thread and 1-thread parsing of the structure fixtures, which
//            must agree; truncated and byte-flipped variants must parse or
//            fail with a status, never crash.
//   tsv      16-thread TSV writing of a random triangle, read back exactly.
//   generate multi-threaded synthetic ensembles, subset-stable.
//   queue    the serving loop's queue, admission, merging and tickets over a
//            host executor (no GPU): 4 client threads submit, wait, poll,
//            cancel and release concurrently; every waited result equals the
//            request's result alone; the ticket table ends empty; a tight
//            budget admits batches one at a time.
//   gpu      (with --gpu) the default context from several threads, and a
//            device server with concurrent submit / release / wait.
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../bioshell_amd/csrc/common.hpp"
#include "../../bioshell_amd/csrc/serve.hpp"
#include "../../include/bioshell_gpu.h"

namespace {

int g_failures = 0;

#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "CHECK failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            ++g_failures;                                                        \
        }                                                                        \
    } while (0)

std::string read_gz(const std::string& path)
{
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) return {};
    std::string out;
    char buf[1 << 16];
    int n;
    while ((n = gzread(f, buf, sizeof buf)) > 0) out.append(buf, (size_t)n);
    gzclose(f);
    return out;
}

struct Parsed {
    bs_status st = BS_OK;
    bs_ensemble_info info{};
    std::vector<double> xyz;
};

Parsed parse(const std::string& text, int selection, int threads)
{
    Parsed p;
    bs_ensemble* e = nullptr;
    p.st = bs_ensemble_parse(text.data(), text.size(), BS_FORMAT_AUTO, selection, nullptr, threads, &e);
    if (p.st != BS_OK) return p;
    bs_ensemble_get_info(e, &p.info);
    p.xyz.resize(p.info.n_models * p.info.n_atoms * 3);
    if (!p.xyz.empty()) bs_ensemble_copy_coords(e, p.xyz.data());
    for (size_t k = 0; k < std::min<size_t>(p.info.n_atoms, 8); ++k) {
        bs_atom_record rec;
        CHECK(bs_ensemble_get_atom(e, k, &rec) == BS_OK);
    }
    bs_ensemble_destroy(e);
    return p;
}

void section_ingest(const std::string& golden)
{
    const char* names[] = {"2jqb.cif.gz", "2gb1.cif.gz", "2gb1.pdb.gz", "2fdo.cif.gz", "1c5n.cif.gz",
                           "5edw.cif.gz", "6ins.cif.gz", "1akd_fragm.pdb.gz", "alphaHelix.pdb.gz"};
    std::mt19937_64 rng(7);
    for (const char* nm : names) {
        const std::string text = read_gz(golden + "/structures/" + nm);
        CHECK(!text.empty());
        for (int sel : {BS_SELECT_ALL, BS_SELECT_CA, BS_SELECT_NOT_HYDROGEN}) {
            const Parsed a = parse(text, sel, 16), b = parse(text, sel, 1);
            CHECK(a.st == b.st);
            CHECK(a.xyz == b.xyz);
        }
        // damaged copies: every outcome is a status, never a crash
        for (int t = 0; t < 24; ++t) {
            std::string bad = text.substr(0, (size_t)(rng() % (text.size() + 1)));
            for (int f = 0; f < 4 && !bad.empty(); ++f) bad[rng() % bad.size()] = (char)(rng() & 0xff);
            const Parsed p = parse(bad, BS_SELECT_ALL, (t & 1) ? 16 : 1);
            CHECK(p.st == BS_OK || p.st == BS_ERR_PARSE || p.st == BS_ERR_INVALID_ARGUMENT);
        }
    }
    // the multi-model fixture by 16 threads: the reference's counts
    const std::string base = read_gz(golden + "/structures/2jqb.cif.gz");
    const Parsed one = parse(base, BS_SELECT_CA, 16);
    CHECK(one.st == BS_OK && one.info.n_models == 20 && one.info.n_atoms == 15);
    std::printf("ok ingest\n");
}

void section_tsv()
{
    const size_t n = 1500;
    std::vector<float> up(n * (n - 1) / 2);
    std::mt19937 rng(3);
    std::uniform_real_distribution<float> U(0.0f, 30.0f);
    for (auto& v : up) v = U(rng);
    // One file per process: the ASan and TSan builds may run at the same time.
    const char* tmp = std::getenv("TMPDIR");
    const std::string file = std::string(tmp && *tmp ? tmp : "/tmp") + "/sanitize_host_" +
                             std::to_string(getpid()) + ".tsv";
    const char* path = file.c_str();
    CHECK(bs_write_distance_tsv(up.data(), n, nullptr, -1, 0, 16, path) == BS_OK);
    std::ifstream in(path);
    std::string a, b;
    double v;
    size_t k = 0;
    bool exact = true;
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) {
            if (!(in >> a >> b >> v)) {
                exact = false;
                break;
            }
            exact = exact && a == std::to_string(i) && b == std::to_string(j) && (float)v == up[k];
            ++k;
        }
    CHECK(exact && k == up.size());
    std::remove(path);
    std::printf("ok tsv\n");
}

void section_generate()
{
    const size_t n = 300, L = 50;
    std::vector<double> all(n * L * 3), part(40 * L * 3);
    CHECK(bs_generate_ensemble(5, n, L, 0.5, 3.0, 0.05, 0.01, all.data()) == BS_OK);
    CHECK(bs_generate_ensemble(5, 40, L, 0.5, 3.0, 0.05, 0.01, part.data()) == BS_OK);
    CHECK(std::equal(part.begin(), part.end(), all.begin()));
    std::printf("ok generate\n");
}

// ------------------------------------------------------------------ queue
// Score of query a against member m: a fixed function of the query's own
// coordinates, so a request's result never depends on its batch.
float host_score(const double* q, int L, int m)
{
    double s = 0.0;
    for (int c = 0; c < 3 * L; ++c) s += q[c] * (1.0 + 0.001 * c);
    return (float)std::fabs(std::fmod(std::fabs(s) * 7.0 + m * 0.37, 50.0));
}

void host_topk(const double* q, int L, int n_lib, int k, uint32_t* idx, float* val)
{
    std::vector<std::pair<float, uint32_t>> v(n_lib);
    for (int m = 0; m < n_lib; ++m) v[m] = {host_score(q, L, m), (uint32_t)m};
    std::partial_sort(v.begin(), v.begin() + k, v.end());
    for (int r = 0; r < k; ++r) {
        idx[r] = v[r].second;
        val[r] = v[r].first;
    }
}

class HostExecutor final : public bs::BatchExecutor {
public:
    explicit HostExecutor(std::vector<bs::ServedLibrary> libs, unsigned seed) : libs_(std::move(libs)), rng_(seed) {}
    double* staging(size_t doubles) override
    {
        q_.resize(doubles);
        return q_.data();
    }
    void launch(size_t lib, int n_q, int k) override
    {
        const int L = libs_[lib].L, n = libs_[lib].n;
        out_.assign((size_t)n_q * k * 2, 0u);
        float* val = reinterpret_cast<float*>(out_.data() + (size_t)n_q * k);
        for (int a = 0; a < n_q; ++a)
            host_topk(q_.data() + (size_t)a * L * 3, L, n, k, out_.data() + (size_t)a * k, val + (size_t)a * k);
        ready_ = std::chrono::steady_clock::now() + std::chrono::microseconds(rng_() % 300);
    }
    bool done() override { return std::chrono::steady_clock::now() >= ready_; }
    const uint32_t* results() const override { return out_.data(); }

private:
    std::vector<bs::ServedLibrary> libs_;
    std::mt19937 rng_;
    std::vector<double> q_;
    std::vector<uint32_t> out_;
    std::chrono::steady_clock::time_point ready_;
};

bs_server* host_server(const std::vector<bs::ServedLibrary>& libs, int slots, size_t budget)
{
    bs::ServerSetup cfg;
    cfg.libs = libs;
    cfg.budget = budget;
    cfg.merge_queries = 24;
    cfg.batch_bytes = [](size_t, int n_q, int k) { return (size_t)n_q * 1000 + (size_t)k; };
    cfg.k_supported = [libs](size_t lib, int k) { return k >= 1 && k <= std::min(64, libs[lib].n); };
    for (int s = 0; s < slots; ++s) cfg.executors.push_back(std::make_unique<HostExecutor>(libs, 100 + s));
    return bs::make_server(std::move(cfg));
}

void section_queue()
{
    const std::vector<bs::ServedLibrary> libs = {{400, 6}, {150, 11}};
    for (size_t budget : {(size_t)0, (size_t)30000}) {
        bs_server* srv = host_server(libs, 3, budget);
        std::atomic<int> bad{0}, waited{0};
        auto client = [&](unsigned seed) {
            std::mt19937 rng(seed);
            for (int r = 0; r < 150; ++r) {
                const size_t lib = rng() % libs.size();
                const int L = libs[lib].L, n_q = 1 + (int)(rng() % 8), k = 1 + (int)(rng() % 40);
                std::vector<double> q((size_t)n_q * L * 3);
                for (auto& v : q) v = (double)(rng() % 10000) * 1e-3;
                std::vector<uint32_t> idx((size_t)n_q * k);
                std::vector<float> val((size_t)n_q * k);
                uint64_t t = 0;
                if (bs_server_submit(srv, lib, q.data(), n_q, k, idx.data(), val.data(), &t) != BS_OK) {
                    ++bad;
                    continue;
                }
                const std::vector<double> qc = q;
                q.assign(q.size(), -1.0);  // the server copied the queries
                const int what = (int)(rng() % 4);
                if (what == 0) {  // drop at once: never written, record goes
                    bs_server_release(srv, t);
                    continue;
                }
                if (what == 1 && bs_server_cancel(srv, t) == BS_OK) {
                    if (bs_server_wait(srv, t) != BS_ERR_CANCELLED) ++bad;
                    bs_server_release(srv, t);
                    continue;
                }
                if (what == 2)
                    while (bs_server_poll(srv, t) == BS_ERR_NOT_READY) std::this_thread::yield();
                if (bs_server_wait(srv, t) != BS_OK) {
                    ++bad;
                } else {
                    ++waited;
                    // the result of the request alone, whatever it was merged with
                    std::vector<uint32_t> wi(k);
                    std::vector<float> wv(k);
                    for (int a = 0; a < n_q; ++a) {
                        host_topk(qc.data() + (size_t)a * L * 3, L, libs[lib].n, k, wi.data(), wv.data());
                        if (!std::equal(wi.begin(), wi.end(), idx.begin() + (size_t)a * k) ||
                            !std::equal(wv.begin(), wv.end(), val.begin() + (size_t)a * k))
                            ++bad;
                    }
                    double ms = 0.0;
                    if (bs_server_latency_ms(srv, t, &ms) != BS_OK || ms < 0.0) ++bad;
                }
                bs_server_release(srv, t);
            }
        };
        std::vector<std::thread> th;
        for (unsigned c = 0; c < 4; ++c) th.emplace_back(client, 11 + c);
        for (auto& t : th) t.join();
        CHECK(bad.load() == 0);
        CHECK(waited.load() > 50);
        // every record was released; running ones drop when they end
        for (int spin = 0; spin < 2000 && bs_server_ticket_count(srv) != 0; ++spin)
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        CHECK(bs_server_ticket_count(srv) == 0);

        // results equal the request alone, exactly, also when merged
        const int L = libs[0].L, n_q = 3, k = 17;
        std::vector<double> q((size_t)n_q * L * 3);
        for (size_t c = 0; c < q.size(); ++c) q[c] = 0.01 * (double)(c % 97);
        std::vector<uint32_t> want((size_t)n_q * k), got((size_t)n_q * k);
        std::vector<float> wantv((size_t)n_q * k), gotv((size_t)n_q * k);
        for (int a = 0; a < n_q; ++a)
            host_topk(q.data() + (size_t)a * L * 3, L, libs[0].n, k, want.data() + (size_t)a * k,
                      wantv.data() + (size_t)a * k);
        std::vector<uint64_t> ts;
        std::vector<std::vector<uint32_t>> gi(6, std::vector<uint32_t>((size_t)n_q * 40));
        std::vector<std::vector<float>> gv(6, std::vector<float>((size_t)n_q * 40));
        uint64_t t0 = 0;
        CHECK(bs_server_submit(srv, 0, q.data(), n_q, k, got.data(), gotv.data(), &t0) == BS_OK);
        for (int o = 0; o < 6; ++o) {  // company with larger k in the same batch
            uint64_t t = 0;
            CHECK(bs_server_submit(srv, 0, q.data(), n_q, 40, gi[o].data(), gv[o].data(), &t) == BS_OK);
            ts.push_back(t);
        }
        CHECK(bs_server_wait(srv, t0) == BS_OK);
        CHECK(got == want && gotv == wantv);
        for (auto t : ts) {
            CHECK(bs_server_wait(srv, t) == BS_OK);
            bs_server_release(srv, t);
        }
        bs_server_release(srv, t0);
        // out of range k fails at submit
        uint64_t tk = 0;
        CHECK(bs_server_submit(srv, 1, q.data(), 1, 65, got.data(), gotv.data(), &tk) == BS_ERR_UNSUPPORTED);
        bs_server_destroy(srv);
    }
    std::printf("ok queue\n");
}

// ------------------------------------------------------------------ gpu
void section_gpu()
{
    if (bs_device_count() < 1) {
        std::printf("skip gpu (no device)\n");
        return;
    }
    const size_t n = 64, L = 40;
    std::vector<double> x(n * L * 3);
    CHECK(bs_generate_ensemble(9, n, L, 0.5, 3.0, 0.0, 0.01, x.data()) == BS_OK);
    std::vector<float> want(n * (n - 1) / 2);
    CHECK(bs_crmsd_all_vs_all(nullptr, x.data(), n, L, BS_BF16X3_GRAM, want.data(), 0) == BS_OK);
    std::atomic<int> bad{0};
    std::vector<std::thread> th;
    for (int t = 0; t < 4; ++t)
        th.emplace_back([&] {
            for (int r = 0; r < 5; ++r) {
                std::vector<float> got(want.size());
                if (bs_crmsd_all_vs_all(nullptr, x.data(), n, L, BS_BF16X3_GRAM, got.data(), 0) != BS_OK ||
                    got != want)
                    ++bad;
                double p = 0.0;
                if (bs_crmsd_pair_f64(nullptr, x.data(), x.data() + L * 3, L, &p) != BS_OK) ++bad;
            }
        });
    for (auto& t : th) t.join();
    CHECK(bad.load() == 0);

    bs_context* ctx = nullptr;
    CHECK(bs_context_create(0, 0, &ctx) == BS_OK);
    bs_library* lib = nullptr;
    const size_t n_lib = 3000, Lq = 32;
    std::vector<double> lx((n_lib + 16) * Lq * 3);
    CHECK(bs_generate_ensemble(4, n_lib + 16, Lq, 0.5, 3.0, 0.05, 0.01, lx.data()) == BS_OK);
    CHECK(bs_library_create(ctx, lx.data(), n_lib, Lq, &lib) == BS_OK);
    const double* q = lx.data() + n_lib * Lq * 3;
    const int k = 20;
    std::vector<uint32_t> wi(16 * k);
    std::vector<float> wv(16 * k);
    CHECK(bs_query_topk(ctx, lib, q, 16, k, wi.data(), wv.data()) == BS_OK);
    bs_server* srv = nullptr;
    CHECK(bs_server_create(ctx, &lib, 1, 0, 3, &srv) == BS_OK);
    th.clear();
    for (int c = 0; c < 4; ++c)
        th.emplace_back([&, c] {
            std::mt19937 rng(c);
            for (int r = 0; r < 60; ++r) {
                const int s = (int)(rng() % 8), m = 1 + (int)(rng() % 8);
                std::vector<uint32_t> gi((size_t)m * k);
                std::vector<float> gv((size_t)m * k);
                uint64_t t = 0;
                if (bs_server_submit(srv, 0, q + (size_t)s * Lq * 3, m, k, gi.data(), gv.data(), &t) != BS_OK) {
                    ++bad;
                    continue;
                }
                if (rng() % 3 == 0) {
                    bs_server_release(srv, t);
                    continue;
                }
                if (bs_server_wait(srv, t) != BS_OK ||
                    !std::equal(gi.begin(), gi.end(), wi.begin() + (size_t)s * k) ||
                    !std::equal(gv.begin(), gv.end(), wv.begin() + (size_t)s * k))
                    ++bad;
                bs_server_release(srv, t);
            }
        });
    for (auto& t : th) t.join();
    CHECK(bad.load() == 0);
    bs_server_destroy(srv);
    bs_library_destroy(lib);
    bs_context_destroy(ctx);
    std::printf("ok gpu\n");
}

}  // namespace

int main(int argc, char** argv)
{
    std::string golden = "tests/golden";
    bool gpu = false;
    for (int a = 1; a < argc; ++a) {
        if (std::strcmp(argv[a], "--gpu") == 0)
            gpu = true;
        else if (std::strncmp(argv[a], "--golden=", 9) == 0)
            golden = argv[a] + 9;
    }
    section_ingest(golden);
    section_tsv();
    section_generate();
    section_queue();
    if (gpu) section_gpu();
    if (g_failures) {
        std::fprintf(stderr, "%d checks failed\n", g_failures);
        return 1;
    }
    std::printf("all sections passed\n");
    return 0;
}
