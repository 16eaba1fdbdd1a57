// tsv.cpp -- a distance matrix written as the three-column TSV that BioShell's
// clust tool reads (bioshell-clustering/src/bin/clust.rs:102-107, through
// DataMatrixBuilder; fixtures tests/test_files/cities.tsv and d5.tsv):
// "label_i<TAB>label_j<TAB>value" per pair, rows in order, i < j (and, on
// request, j > i too).
//
// Values are the f32 distances: by default in the shortest positional form
// that reads back to the same f32 (what Rust's `{}` prints for an f32), or
// with a fixed number of decimals of the exact f32 value, rounded to nearest.
// Rows are formatted by up to 16 threads in bounded rounds and written in
// order, so a 10k-element matrix (50M lines) never sits in memory as text.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

namespace bs {
namespace {

inline size_t tri(size_t n, size_t i, size_t j) { return i * n - i * (i + 1) / 2 + (j - i - 1); }

void put_value(std::string& out, float v, int decimals)
{
    char buf[128];
    if (std::isnan(v)) {
        out += "NaN";
        return;
    }
    if (std::isinf(v)) {
        out += v < 0 ? "-inf" : "inf";
        return;
    }
    if (decimals < 0) {
        // the shortest digits that read back to v (scientific to_chars),
        // placed positionally with zeros as needed, as Rust's `{}` does
        const auto r = std::to_chars(buf, buf + sizeof buf - 1, v, std::chars_format::scientific);
        *r.ptr = '\0';  // to_chars does not terminate; atoi below reads the exponent
        const char* p = buf;
        if (*p == '-') out += *p++;
        std::string digits;
        while (p < r.ptr && *p != 'e') {
            if (*p != '.') digits += *p;
            ++p;
        }
        const int exp = p < r.ptr ? std::atoi(p + 1) : 0;
        if (exp >= 0) {
            const size_t int_len = (size_t)exp + 1;
            if (digits.size() <= int_len) {
                out += digits;
                out.append(int_len - digits.size(), '0');
            } else {
                out.append(digits, 0, int_len);
                out += '.';
                out.append(digits, int_len, std::string::npos);
            }
        } else {
            out += "0.";
            out.append((size_t)(-exp - 1), '0');
            out += digits;
        }
    } else {
        const int k = std::snprintf(buf, sizeof buf, "%.*f", decimals, (double)v);
        out.append(buf, (size_t)k);
    }
}

}  // namespace
}  // namespace bs

using namespace bs;

extern "C" bs_status bs_write_distance_tsv(const float* upper, size_t n, const char* const* labels,
                                           int decimals, int both_triangles, int n_threads,
                                           const char* path)
{
    return guard([&] {
        BS_REQUIRE(path, "null path");
        BS_REQUIRE(n < 2 || upper, "null matrix");
        BS_REQUIRE(decimals >= -1 && decimals <= 30, "decimals must be -1 (shortest) or 0..30, got %d", decimals);
        std::vector<std::string> names(n);
        for (size_t i = 0; i < n; ++i) {
            if (labels) {
                BS_REQUIRE(labels[i], "label %zu is null", i);
                names[i] = labels[i];
            } else {
                names[i] = std::to_string(i);
            }
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) fail(BS_ERR_IO, "cannot open %s for writing", path);
        std::vector<char> iobuf(1 << 22);
        std::setvbuf(f, iobuf.data(), _IOFBF, iobuf.size());
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int nt = std::max(1, n_threads > 0 ? std::min(n_threads, 64) : (int)std::min(16u, hw));
        const size_t per_row_full = n > 0 ? n - 1 : 0;
        auto lines_of = [&](size_t i) { return both_triangles ? per_row_full : n - 1 - i; };
        const size_t kLinesPerTask = 1 << 18;
        bool ok = true;
        size_t row = 0;
        while (row < n && ok) {
            // cut the next nt tasks of ~kLinesPerTask lines each at row boundaries
            std::vector<std::pair<size_t, size_t>> tasks;
            while ((int)tasks.size() < nt && row < n) {
                size_t r1 = row, lines = 0;
                while (r1 < n && (lines == 0 || lines + lines_of(r1) <= kLinesPerTask)) lines += lines_of(r1++);
                tasks.push_back({row, r1});
                row = r1;
            }
            std::vector<std::string> text(tasks.size());
            auto work = [&](size_t t) {
                std::string& s = text[t];
                for (size_t i = tasks[t].first; i < tasks[t].second; ++i) {
                    const size_t j0 = both_triangles ? 0 : i + 1;
                    for (size_t j = j0; j < n; ++j) {
                        if (j == i) continue;
                        const float v = i < j ? upper[tri(n, i, j)] : upper[tri(n, j, i)];
                        s += names[i];
                        s += '\t';
                        s += names[j];
                        s += '\t';
                        put_value(s, v, decimals);
                        s += '\n';
                    }
                }
            };
            std::vector<std::thread> th;
            for (size_t t = 1; t < tasks.size(); ++t) th.emplace_back(work, t);
            work(0);
            for (auto& x : th) x.join();
            for (const std::string& s : text)
                if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) ok = false;
        }
        const int rc = std::fclose(f);
        if (!ok || rc != 0) fail(BS_ERR_IO, "cannot write %s", path);
    });
}
