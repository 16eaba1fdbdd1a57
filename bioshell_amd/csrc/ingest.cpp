// ingest.cpp -- multi-model mmCIF / PDB text -> conformational ensemble: the
// atom records of the first model and the coordinates of the selected atoms
// in every model, laid out n_models x n_atoms x 3 (f64) ready for upload.
//
// The result is what BioShell's own reading gives: Deposit::from_file
// (bioshell-pdb/src/deposit.rs:81-85) builds a Structure whose atoms come from
// the first model and whose model_coordinates hold every model; the caller
// sets each model (structure.rs:277-285) and keeps the atoms an atom predicate
// accepts.  Followed here:
//   mmCIF  bioshell-cif/src/cif_line_iterator.rs:16-40 (';' text fields),
//          bioshell-cif/src/lib.rs:596-627 (line kinds), 653-760 (blocks and
//          loops), 220-247 + 331-337 (loop rows), 499-501 (get_loop);
//          bioshell-core/src/io/split_into_strings.rs:15-36, 70-127 (tokens,
//          quotes kept); bioshell-cif/src/column_mapping.rs:101-190;
//          bioshell-pdb/src/pdb_atom.rs:130-203 (atoms, models), 306-331
//          (atom names)
//   PDB    bioshell-pdb/src/load_pdb.rs:43-115, pdb_atom.rs:87-116,
//          bioshell-core/src/vec3.rs:241-246
//   format load_cif.rs:155-167, load_pdb.rs:225-237
//   counts structure.rs:236-255; predicates pdb_atom_filters.rs:191-199,
//          351-358, 388-422, pdb_atom.rs:352-365
//
// Work split: the text is cut into chunks at line boundaries and scanned by up
// to 16 threads.  The CIF state machine runs sequentially over compact
// per-chunk event lists (one event per header line or per run of data lines),
// the PDB model bookkeeping over per-chunk MODEL/ENDMDL events, and the
// per-atom parsing is parallel again.  A CIF loop whose rows do not coincide
// with lines (or that holds ';' text fields) is read by a sequential path.
#include <fcntl.h>
#include <locale.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <memory>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "ingest.hpp"

namespace bs {
namespace {

using u8 = unsigned char;

[[noreturn]] void parse_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void parse_error(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(BS_ERR_PARSE, buf);
}

// ------------------------------------------------------------ characters

struct AsciiClass {
    bool ws[128];
    AsciiClass()
    {
        for (int c = 0; c < 128; ++c) ws[c] = c == 0x20 || (c >= 0x09 && c <= 0x0d);
    }
};
const AsciiClass kAscii;

// Byte length of the whitespace character (char::is_whitespace) at p, or 0.
inline int ws_at(const u8* p, const u8* e)
{
    const u8 c = *p;
    if (c < 0x80) return kAscii.ws[c] ? 1 : 0;
    const ptrdiff_t n = e - p;
    if (c == 0xC2) return (n >= 2 && (p[1] == 0x85 || p[1] == 0xA0)) ? 2 : 0;
    if (n < 3) return 0;
    if (c == 0xE1) return (p[1] == 0x9A && p[2] == 0x80) ? 3 : 0;
    if (c == 0xE2) {
        if (p[1] == 0x80) return (p[2] <= 0x8A || p[2] == 0xA8 || p[2] == 0xA9 || p[2] == 0xAF) ? 3 : 0;
        if (p[1] == 0x81) return p[2] == 0x9F ? 3 : 0;
        return 0;
    }
    if (c == 0xE3) return (p[1] == 0x80 && p[2] == 0x80) ? 3 : 0;
    return 0;
}

inline const u8* last_char(const u8* b, const u8* e)
{
    const u8* p = e - 1;
    while (p > b && (*p & 0xC0) == 0x80) --p;
    return p;
}

inline void trim(const u8*& b, const u8*& e)
{
    while (b < e) {
        const int k = ws_at(b, e);
        if (!k) break;
        b += k;
    }
    while (e > b) {
        const u8* p = last_char(b, e);
        if (ws_at(p, e) != e - p) break;
        e = p;
    }
}

inline const u8* skip_ws(const u8* p, const u8* e)
{
    while (p < e) {
        const int k = ws_at(p, e);
        if (!k) break;
        p += k;
    }
    return p;
}

// Continuation bytes never start a whitespace character, so stepping bytewise
// through non-whitespace is exact.
inline const u8* skip_nonws(const u8* p, const u8* e)
{
    while (p < e && !ws_at(p, e)) ++p;
    return p;
}

inline bool starts(const u8* b, const u8* e, const char* w)
{
    const size_t n = std::strlen(w);
    return (size_t)(e - b) >= n && std::memcmp(b, w, n) == 0;
}

// is_quote_char (split_into_strings.rs:127): " ' and the typographic quotes;
// returns a code identifying the character, 0 if it is not a quote.
inline int quote_first(const u8* b, const u8* e)
{
    if (*b == '"') return 1;
    if (*b == '\'') return 2;
    if (*b == 0xE2 && e - b >= 3 && b[1] == 0x80) {
        switch (b[2]) {
        case 0x9C: return 3;
        case 0x9D: return 4;
        case 0x98: return 5;
        case 0x99: return 6;
        }
    }
    return 0;
}

inline int quote_last(const u8* b, const u8* e)
{
    const u8 c = e[-1];
    if (c == '"') return 1;
    if (c == '\'') return 2;
    if (e - b >= 3 && e[-3] == 0xE2 && e[-2] == 0x80) {
        switch (c) {
        case 0x9C: return 3;
        case 0x9D: return 4;
        case 0x98: return 5;
        case 0x99: return 6;
        }
    }
    return 0;
}

// quote_style (split_into_strings.rs:106-125) == Begin for a token that does
// not start with whitespace.
inline bool quote_open(const u8* b, const u8* e)
{
    const int f = quote_first(b, e);
    if (!f) return false;
    if (e - b == 1) return true;  // a lone ASCII quote (str::len counts bytes)
    return quote_last(b, e) != f;
}

// split_into_strings(line, false) over a trimmed, non-empty line: calls
// emit(b, e) for each token and returns the count.  Tokens are contiguous
// spans of the line: a token that opens a quote absorbs the following words,
// whitespace included, until it ends with the same quote.
template <class F>
inline size_t tokenize(const u8* b, const u8* e, F&& emit)
{
    size_t n = 0;
    const u8* tb = b;
    const u8* te = skip_nonws(b, e);
    const u8* p = te;
    while (p < e) {
        const u8* q = skip_ws(p, e);
        const u8* r = skip_nonws(q, e);
        if (quote_open(tb, te)) {
            te = r;
        } else {
            emit(tb, te);
            ++n;
            tb = q;
            te = r;
        }
        p = r;
    }
    emit(tb, te);
    return n + 1;
}

// ------------------------------------------------------------ numbers

locale_t c_locale()
{
    static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
    return loc;
}

const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                           1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

inline bool is_digit(u8 c) { return c >= '0' && c <= '9'; }

// <f64 as FromStr>: [+-] (digits [. digits*] | . digits) [(e|E) [+-] digits]
// or inf / infinity / nan in any case; correctly rounded.  Mantissas below
// 2^53 with |exponent| <= 22 take one exact IEEE operation (Clinger's fast
// path), everything else strtod in the C locale.
bool parse_f64(const u8* b, const u8* e, double* out)
{
    const u8* p = b;
    bool neg = false;
    if (p < e && (*p == '+' || *p == '-')) {
        neg = *p == '-';
        ++p;
    }
    if (p == e) return false;
    auto word = [&](const char* w) {
        const size_t n = std::strlen(w);
        if ((size_t)(e - p) != n) return false;
        for (size_t i = 0; i < n; ++i)
            if ((p[i] | 0x20) != (u8)w[i]) return false;
        return true;
    };
    if (word("inf") || word("infinity")) {
        *out = neg ? -HUGE_VAL : HUGE_VAL;
        return true;
    }
    if (word("nan")) {
        *out = std::numeric_limits<double>::quiet_NaN();
        return true;
    }
    // digits beyond 17 significant ones make `big`: strtod then decides
    uint64_t mant = 0;
    bool big = false;
    int64_t dexp = 0;
    const u8* q = p;
    while (p < e && is_digit(*p)) {
        if (mant <= 9999999999999999ULL)
            mant = mant * 10 + (uint64_t)(*p - '0');
        else
            big = true;
        ++p;
    }
    const bool int_digits = p > q;
    bool frac_digits = false;
    if (p < e && *p == '.') {
        ++p;
        const u8* f = p;
        while (p < e && is_digit(*p)) {
            if (mant <= 9999999999999999ULL) {
                mant = mant * 10 + (uint64_t)(*p - '0');
                --dexp;
            } else {
                big = true;
            }
            ++p;
        }
        frac_digits = p > f;
    }
    if (!int_digits && !frac_digits) return false;
    if (p < e && (*p == 'e' || *p == 'E')) {
        ++p;
        bool eneg = false;
        if (p < e && (*p == '+' || *p == '-')) {
            eneg = *p == '-';
            ++p;
        }
        if (p == e || !is_digit(*p)) return false;
        int64_t ex = 0;
        while (p < e && is_digit(*p)) {
            if (ex < 1000000) ex = ex * 10 + (*p - '0');
            ++p;
        }
        dexp += eneg ? -ex : ex;
    }
    if (p != e) return false;
    if (!big && mant <= (1ULL << 53) && dexp >= -22 && dexp <= 22) {
        double v = (double)mant;
        v = dexp < 0 ? v / kPow10[-dexp] : v * kPow10[dexp];
        *out = neg ? -v : v;
        return true;
    }
    std::string s(reinterpret_cast<const char*>(b), (size_t)(e - b));
    char* end = nullptr;
    *out = strtod_l(s.c_str(), &end, c_locale());
    return true;
}

bool parse_i32(const u8* b, const u8* e, int32_t* out)
{
    bool neg = false;
    if (b < e && (*b == '+' || *b == '-')) {
        neg = *b == '-';
        ++b;
    }
    if (b == e) return false;
    int64_t v = 0;
    for (; b < e; ++b) {
        if (!is_digit(*b)) return false;
        v = v * 10 + (*b - '0');
        if (v > 2147483648LL) return false;
    }
    if (neg) v = -v;
    if (v > 2147483647LL) return false;
    *out = (int32_t)v;
    return true;
}

bool parse_usize(const u8* b, const u8* e, uint64_t* out)
{
    if (b < e && *b == '+') ++b;
    if (b == e) return false;
    uint64_t v = 0;
    for (; b < e; ++b) {
        if (!is_digit(*b)) return false;
        const uint64_t d = (uint64_t)(*b - '0');
        if (v > (std::numeric_limits<uint64_t>::max() - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

std::string str(const u8* b, const u8* e) { return std::string(reinterpret_cast<const char*>(b), (size_t)(e - b)); }

// ------------------------------------------------------------ text layout

struct Text {
    const u8* t;
    size_t n;
};

struct Chunk {
    size_t b, e;
};

std::vector<Chunk> make_chunks(const Text& x, size_t want)
{
    std::vector<Chunk> out;
    size_t pos = 0;
    for (size_t k = 1; k <= want && pos < x.n; ++k) {
        size_t cut = x.n;
        if (k < want) {
            const size_t target = x.n * k / want;
            if (target < pos) continue;
            const void* nl = std::memchr(x.t + target, '\n', x.n - target);
            cut = nl ? (size_t)(static_cast<const u8*>(nl) - x.t) + 1 : x.n;
        }
        if (cut > pos) {
            out.push_back({pos, cut});
            pos = cut;
        }
    }
    return out;
}

// BufRead::lines over [b, e) of the text: f(line_begin, line_end) without the
// '\n' and a '\r' just before it.
template <class F>
inline void for_lines(const Text& x, size_t b, size_t e, F&& f)
{
    size_t p = b;
    while (p < e) {
        const void* nl = std::memchr(x.t + p, '\n', e - p);
        size_t le = nl ? (size_t)(static_cast<const u8*>(nl) - x.t) : e;
        const size_t next = nl ? le + 1 : e;
        if (nl && le > p && x.t[le - 1] == '\r') --le;
        if (!f(p, le)) return;
        p = next;
    }
}

int thread_count(int requested, size_t bytes)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int nt = requested > 0 ? requested : (int)std::min(16u, hw);
    nt = std::min(nt, 64);
    const size_t by_size = std::max<size_t>(1, bytes >> 20);  // >= 1 MiB per thread
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)nt, by_size));
}

// Runs f(i) for i in [0, n) on nt threads; the first exception is rethrown.
template <class F>
void parallel_for(size_t n, int nt, F&& f)
{
    if (nt <= 1 || n <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&] {
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n) return;
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                if (!err) err = std::current_exception();
                next.store(n);
                return;
            }
        }
    };
    std::vector<std::thread> th;
    const int k = (int)std::min<size_t>((size_t)nt, n);
    for (int t = 1; t < k; ++t) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

// Offset of the first line that is not UTF-8, or x.n.
size_t first_invalid_line(const Text& x)
{
    size_t i = 0;
    while (i < x.n && x.t[i] < 0x80) ++i;
    if (i == x.n) return x.n;
    size_t line = i;
    while (line > 0 && x.t[line - 1] != '\n') --line;
    while (i < x.n) {
        const u8 c = x.t[i];
        if (c == '\n') {
            line = ++i;
            continue;
        }
        if (c < 0x80) {
            ++i;
            continue;
        }
        int len = 0;
        uint32_t cp = 0;
        if (c >= 0xC2 && c <= 0xDF) {
            len = 2;
            cp = c & 0x1F;
        } else if (c >= 0xE0 && c <= 0xEF) {
            len = 3;
            cp = c & 0x0F;
        } else if (c >= 0xF0 && c <= 0xF4) {
            len = 4;
            cp = c & 0x07;
        } else {
            return line;
        }
        if (i + (size_t)len > x.n) return line;
        for (int k = 1; k < len; ++k) {
            if ((x.t[i + k] & 0xC0) != 0x80) return line;
            cp = (cp << 6) | (x.t[i + k] & 0x3F);
        }
        if ((len == 3 && (cp < 0x800 || (cp >= 0xD800 && cp <= 0xDFFF))) ||
            (len == 4 && (cp < 0x10000 || cp > 0x10FFFF)))
            return line;
        i += (size_t)len;
    }
    return x.n;
}

// is_cif_file / is_pdb_file (load_cif.rs:155-167, load_pdb.rs:225-237) on the
// first non-empty line, in Deposit::from_file's order.
int detect_format(const Text& x)
{
    const size_t bad = first_invalid_line(x);
    int fmt = -1;
    for_lines(x, 0, x.n, [&](size_t b, size_t e) {
        if (b >= bad) parse_error("stream did not contain valid UTF-8");
        if (e == b) return true;
        const u8* p = x.t + b;
        const u8* q = x.t + e;
        if (starts(p, q, "data_"))
            fmt = BS_FORMAT_CIF;
        else if (starts(p, q, "HEADER") || starts(p, q, "ATOM") || starts(p, q, "HETATM") ||
                 starts(p, q, "REMARK") || starts(p, q, "MODEL"))
            fmt = BS_FORMAT_PDB;
        return false;
    });
    if (fmt < 0) parse_error("invalid file format: neither mmCIF nor PDB");
    return fmt;
}

// ------------------------------------------------------------ structure

// The intermediate result of either reader: per-atom records of the first
// model (all atoms) and, for every row / ATOM line, its model and position.
struct Parsed {
    std::vector<AtomRecord> atoms;
    size_t n_models = 0;
};

// format_atom_name (pdb_atom.rs:306-331).
std::string format_atom_name(const u8* nb, const u8* ne, const u8* eb, const u8* ee)
{
    trim(nb, ne);
    trim(eb, ee);
    if (nb == ne) parse_error("empty atom name");
    std::string res = "    ";
    auto take_chars = [&](int start, int count) {
        // replaces characters [start, start + count) of res with the name's first chars
        std::string head;
        const u8* p = nb;
        int k = 0;
        while (p < ne && k < count) {
            const u8* q = p + 1;
            while (q < ne && (*q & 0xC0) == 0x80) ++q;
            head.append(reinterpret_cast<const char*>(p), (size_t)(q - p));
            p = q;
            ++k;
        }
        return std::string(res.begin(), res.begin() + start) + head + std::string((size_t)(4 - start - k), ' ');
    };
    if (is_digit(*nb)) return take_chars(0, 4);
    const size_t elen = eb == ee ? 1 : std::min<size_t>((size_t)(ee - eb), 2);
    const int pad = elen == 1 ? 1 : 0;
    return take_chars(pad, 4 - pad);
}

// value_or_default::<char> (bioshell-cif/src/lib.rs:799-801).
std::string one_char(const u8* b, const u8* e, const char* field)
{
    if ((e - b == 1) && (*b == '?' || *b == '.')) return " ";
    const u8* q = b < e ? b + 1 : e;
    while (q < e && (*q & 0xC0) == 0x80) ++q;
    if (b == e || q != e) parse_error("%s is not a single character: '%s'", field, str(b, e).c_str());
    return str(b, e);
}

bool has_value(const u8* b, const u8* e) { return !((e - b == 1) && (*b == '?' || *b == '.')); }

struct Tok {
    const u8* b;
    const u8* e;
};

constexpr int kCols = 16;
// The _atom_site columns PdbAtom::from_cif_data asks for, in its order.
const char* const kAtomSiteCols[kCols] = {
    "id",      "label_atom_id", "label_alt_id", "label_comp_id",  "auth_asym_id", "label_seq_id",
    "pdbx_PDB_ins_code", "Cartn_x", "Cartn_y", "Cartn_z", "occupancy", "B_iso_or_equiv",
    "type_symbol", "pdbx_PDB_model_num", "auth_seq_id", "label_entity_id"};

// create_pdb_atom (pdb_atom.rs:175-203).
AtomRecord cif_atom(const Tok* t)
{
    AtomRecord a;
    if (!parse_i32(t[0].b, t[0].e, &a.serial)) parse_error("invalid atom id '%s'", str(t[0].b, t[0].e).c_str());
    a.element = str(t[12].b, t[12].e);
    a.has_element = true;
    a.name = format_atom_name(t[1].b, t[1].e, t[12].b, t[12].e);
    a.alt_loc = one_char(t[2].b, t[2].e, "label_alt_id");
    a.res_name = str(t[3].b, t[3].e);
    a.chain_id = str(t[4].b, t[4].e);
    const Tok& rs = has_value(t[14].b, t[14].e) ? t[14] : t[5];
    if (!parse_i32(rs.b, rs.e, &a.res_seq)) parse_error("invalid residue number '%s'", str(rs.b, rs.e).c_str());
    a.i_code = one_char(t[6].b, t[6].e, "pdbx_PDB_ins_code");
    if (!parse_f64(t[10].b, t[10].e, &a.occupancy)) parse_error("invalid occupancy '%s'", str(t[10].b, t[10].e).c_str());
    if (!parse_f64(t[11].b, t[11].e, &a.temp_factor))
        parse_error("invalid B factor '%s'", str(t[11].b, t[11].e).c_str());
    a.hetero = false;
    return a;
}

// ------------------------------------------------------------ selection

// format_name! (pdb_atom_filters.rs:351-358): four bytes as they are, else
// centred to four characters.
std::string padded_name(const std::string& n)
{
    if (n.size() == 4) return n;
    size_t chars = 0;
    for (unsigned char c : n) chars += (c & 0xC0) != 0x80;
    if (chars >= 4) return n;
    const size_t left = (4 - chars) / 2;
    return std::string(left, ' ') + n + std::string(4 - chars - left, ' ');
}

// is_hydrogen (pdb_atom.rs:352-365).
bool hydrogen(const AtomRecord& a)
{
    if (a.has_element && a.element == "H") return true;
    const std::string& b = a.name;
    if (b.empty()) return false;
    const unsigned char c = (unsigned char)b[0];
    if (c == 'H') return true;
    if (is_digit(c) || c == ' ' || c == '\t' || c == '\n' || c == '\x0c' || c == '\r')
        return b.size() > 1 && b[1] == 'H';
    return false;
}

bool selected(const AtomRecord& a, int selection, const std::string* chain)
{
    if (chain && a.chain_id != *chain) return false;
    switch (selection) {
    case BS_SELECT_ALL: return true;
    case BS_SELECT_CA: return padded_name(a.name) == " CA ";
    case BS_SELECT_CB: return padded_name(a.name) == " CB ";
    case BS_SELECT_NOT_HYDROGEN: return !hydrogen(a);
    }
    return false;
}

// Per-row model bookkeeping shared by both readers: rows are numbered in file
// order; models[r] is overwritten with (model index << 32 | position).
struct Rows {
    std::vector<uint64_t> v;
    size_t first_phase = 0;  // rows read while only one model had been seen
    std::vector<size_t> count;
};

void finish_structure(Parsed& ps, bs_ensemble& out, int selection, const std::string* chain,
                      std::vector<int32_t>& sel)
{
    out.n_structure_atoms = ps.atoms.size();
    // count_chains / count_residues (structure.rs:236-252)
    {
        std::vector<const std::string*> chains;
        for (const AtomRecord& a : ps.atoms) chains.push_back(&a.chain_id);
        std::sort(chains.begin(), chains.end(), [](const std::string* x, const std::string* y) { return *x < *y; });
        out.n_chains = (size_t)(std::unique(chains.begin(), chains.end(),
                                            [](const std::string* x, const std::string* y) { return *x == *y; }) -
                                chains.begin());
        size_t breaks = 0;
        for (size_t k = 1; k < ps.atoms.size(); ++k) {
            const AtomRecord& x = ps.atoms[k - 1];
            const AtomRecord& y = ps.atoms[k];
            if (!(x.chain_id == y.chain_id && x.res_seq == y.res_seq && x.i_code == y.i_code)) ++breaks;
        }
        out.n_residues = breaks + 1;
    }
    sel.assign(ps.atoms.size(), -1);
    for (size_t k = 0; k < ps.atoms.size(); ++k)
        if (selected(ps.atoms[k], selection, chain)) {
            sel[k] = (int32_t)out.atoms.size();
            out.atoms.push_back(ps.atoms[k]);
        }
}

// set_model for every model (structure.rs:277-285).
void check_models(const std::vector<size_t>& count, size_t n_atoms)
{
    for (size_t m = 0; m < count.size(); ++m)
        if (count[m] != n_atoms)
            parse_error("wrong number of atoms in model %zu: %zu, the first model has %zu", m, count[m], n_atoms);
}

// ------------------------------------------------------------ mmCIF

enum EvKind : uint8_t { EV_BLOCK, EV_NAME, EV_ITEM, EV_LOOP, EV_TEXT, EV_RUN, EV_BAD };

struct Event {
    EvKind kind;
    size_t b = 0, e = 0;          // RUN: first line start .. last line end; TEXT: value span
    size_t k0 = 0, k1 = 0;        // NAME / ITEM key span
    size_t v0 = 0, v1 = 0;        // ITEM value span
    size_t lines = 0, tokens = 0;  // RUN
    size_t first_tokens = 0, min_tokens = 0, max_tokens = 0;
    size_t first_b = 0, first_e = 0;  // RUN: first line's trimmed span
    bool multiline = false;            // TEXT: a ';' field over several lines
};

struct TextField {
    size_t open, close_next;  // open line start; offset after the closing line
    bool closed;
};

// The value the reader stores for a ';' text field: the joined lines trimmed
// (cif_line_iterator.rs:16-40 then lib.rs:596-627).
std::string text_value(const Text& x, const Event& ev)
{
    if (!ev.multiline) return str(x.t + ev.b, x.t + ev.e);
    std::string s;
    for_lines(x, ev.b, ev.e, [&](size_t b, size_t e) {
        s.append(reinterpret_cast<const char*>(x.t + b), e - b);
        s.push_back('\n');
        return true;
    });
    // the closing line carries no '\n' of its own; an unclosed field keeps its last one
    if (ev.v0 == 1 && !s.empty()) s.pop_back();
    const u8* b = reinterpret_cast<const u8*>(s.data());
    const u8* e = b + s.size();
    trim(b, e);
    return std::string(reinterpret_cast<const char*>(b), (size_t)(e - b));
}

struct CifScan {
    std::vector<std::vector<Event>> events;  // per chunk
};

void cif_parse(const Text& x, int nt, int selection, const std::string* chain, bs_ensemble& out)
{
    const std::vector<Chunk> chunks = make_chunks(x, (size_t)nt * 4);
    const size_t nc = chunks.size();
    // 1. lines that open or close ';' text fields
    std::vector<std::vector<size_t>> semis(nc);
    parallel_for(nc, nt, [&](size_t c) {
        for_lines(x, chunks[c].b, chunks[c].e, [&](size_t b, size_t e) {
            if (e > b && x.t[b] == ';') semis[c].push_back(b);
            return true;
        });
    });
    std::vector<TextField> fields;
    {
        bool open = false;
        for (size_t c = 0; c < nc; ++c)
            for (size_t off : semis[c]) {
                if (!open) {
                    fields.push_back({off, x.n, false});
                    open = true;
                } else {
                    const void* nl = std::memchr(x.t + off, '\n', x.n - off);
                    fields.back().close_next = nl ? (size_t)(static_cast<const u8*>(nl) - x.t) + 1 : x.n;
                    fields.back().closed = true;
                    open = false;
                }
            }
    }
    // 2. per-chunk events
    std::vector<std::vector<Event>> events(nc);
    parallel_for(nc, nt, [&](size_t c) {
        std::vector<Event>& ev = events[c];
        size_t fi = (size_t)(std::lower_bound(fields.begin(), fields.end(), chunks[c].b,
                                              [](const TextField& f, size_t off) { return f.close_next <= off; }) -
                             fields.begin());
        for_lines(x, chunks[c].b, chunks[c].e, [&](size_t lb, size_t le) {
            while (fi < fields.size() && fields[fi].close_next <= lb) ++fi;
            if (fi < fields.size() && fields[fi].open <= lb) {
                if (fields[fi].open == lb) {
                    Event t{EV_TEXT};
                    t.b = lb;
                    t.e = fields[fi].close_next;
                    t.multiline = true;
                    t.v0 = fields[fi].closed ? 1 : 0;
                    ev.push_back(t);
                }
                return true;  // inside a text field
            }
            const u8* b = x.t + lb;
            const u8* e = x.t + le;
            trim(b, e);
            if (b == e || *b == '#') return true;
            if (starts(b, e, "data_")) {
                Event t{EV_BLOCK};
                t.b = (size_t)(b - x.t) + 5;
                t.e = (size_t)(e - x.t);
                ev.push_back(t);
            } else if (*b == '_') {
                Event t{EV_NAME};
                size_t n = 0;
                tokenize(b, e, [&](const u8* tb, const u8* te) {
                    if (n == 0) {
                        t.k0 = (size_t)(tb - x.t);
                        t.k1 = (size_t)(te - x.t);
                    } else if (n == 1) {
                        t.v0 = (size_t)(tb - x.t);
                        t.v1 = (size_t)(te - x.t);
                    }
                    ++n;
                });
                t.kind = n == 1 ? EV_NAME : n == 2 ? EV_ITEM : EV_BAD;
                t.b = lb;
                t.e = le;
                ev.push_back(t);
            } else if (starts(b, e, "loop_")) {
                ev.push_back(Event{EV_LOOP});
            } else if (*b == ';') {
                Event t{EV_TEXT};
                t.b = (size_t)(b - x.t);
                t.e = (size_t)(e - x.t);
                ev.push_back(t);
            } else {
                const size_t n = tokenize(b, e, [](const u8*, const u8*) {});
                if (ev.empty() || ev.back().kind != EV_RUN) {
                    Event t{EV_RUN};
                    t.b = lb;
                    t.first_tokens = t.min_tokens = t.max_tokens = n;
                    t.first_b = (size_t)(b - x.t);
                    t.first_e = (size_t)(e - x.t);
                    ev.push_back(t);
                }
                Event& r = ev.back();
                r.e = le;
                r.lines += 1;
                r.tokens += n;
                r.min_tokens = std::min(r.min_tokens, n);
                r.max_tokens = std::max(r.max_tokens, n);
            }
            return true;
        });
    });
    // 3. the block / loop state machine (lib.rs:653-760) over the events
    struct LoopInfo {
        std::vector<std::string> cols;
        size_t tokens = 0;
        std::vector<const Event*> parts;  // RUN and TEXT events, in order
        bool atom_site = false;
    };
    int n_blocks = 0;
    bool in_loop = false;
    LoopInfo cur;
    bool have_atom_site = false;
    LoopInfo atom_site;
    bool item_open = false;
    std::string open_key;
    std::unordered_map<std::string, std::string> items;  // _atom_site.* items of the block
    auto close_loop = [&] {
        if (n_blocks == 0) parse_error("a loop_ outside any data block");
        if (!have_atom_site) {
            for (const std::string& c : cur.cols)
                if (c.compare(0, 11, "_atom_site.") == 0) {
                    have_atom_site = true;
                    atom_site = cur;
                    break;
                }
        }
        in_loop = false;
        cur = LoopInfo{};
    };
    auto set_item = [&](const std::string& k, std::string v) {
        if (n_blocks == 0) parse_error("a data item outside any data block");
        if (k.compare(0, 11, "_atom_site.") == 0) items[k] = std::move(v);
    };
    for (size_t c = 0; c < nc; ++c)
        for (const Event& ev : events[c]) {
            switch (ev.kind) {
            case EV_BAD: parse_error("invalid CIF line: %s", str(x.t + ev.b, x.t + ev.e).c_str());
            case EV_LOOP:
                if (in_loop) close_loop();
                in_loop = true;
                cur = LoopInfo{};
                break;
            case EV_BLOCK:
                if (in_loop) close_loop();
                ++n_blocks;
                break;
            case EV_ITEM:
                if (in_loop) close_loop();
                set_item(str(x.t + ev.k0, x.t + ev.k1), str(x.t + ev.v0, x.t + ev.v1));
                break;
            case EV_TEXT:
                if (in_loop) {
                    cur.tokens += 1;
                    cur.parts.push_back(&ev);
                } else if (item_open) {
                    set_item(open_key, text_value(x, ev));
                    item_open = false;
                } else {
                    parse_error("a multi-line string outside a data item");
                }
                break;
            case EV_NAME: {
                std::string name = str(x.t + ev.k0, x.t + ev.k1);
                if (in_loop) {
                    if (cur.tokens == 0) {
                        cur.cols.push_back(std::move(name));
                        break;
                    }
                    close_loop();
                }
                if (item_open) parse_error("dangling data item %s", name.c_str());
                item_open = true;
                open_key = std::move(name);
                break;
            }
            case EV_RUN:
                if (in_loop) {
                    cur.tokens += ev.tokens;
                    cur.parts.push_back(&ev);
                } else if (item_open && ev.lines == 1 && ev.first_tokens == 1) {
                    set_item(open_key, str(x.t + ev.first_b, x.t + ev.first_e));
                    item_open = false;
                } else {
                    parse_error("data values outside a loop");
                }
                break;
            }
        }
    if (in_loop) close_loop();
    if (n_blocks > 1) parse_error("more than one data block");
    if (n_blocks == 0) parse_error("no data block");

    // 4. _atom_site rows (column_mapping.rs:101-190)
    int col[kCols];
    size_t n_cols = 0;
    std::vector<std::vector<std::string>> slow_rows;  // the 16 values per row, sequential path
    bool fast = false;
    size_t n_rows = 0;
    if (have_atom_site) {
        n_cols = atom_site.cols.size();
        for (int k = 0; k < kCols; ++k) {
            col[k] = -1;
            for (size_t j = 0; j < n_cols; ++j)
                if (atom_site.cols[j].find(kAtomSiteCols[k]) != std::string::npos) {
                    col[k] = (int)j;
                    break;
                }
            if (col[k] < 0) parse_error("missing CIF loop key %s", kAtomSiteCols[k]);
        }
        fast = true;
        for (const Event* p : atom_site.parts)
            if (p->kind != EV_RUN || p->min_tokens != n_cols || p->max_tokens != n_cols) fast = false;
        if (fast) {
            n_rows = atom_site.tokens / n_cols;
        } else {
            // rows are the loop's token stream cut every n_cols tokens
            std::vector<std::string> row;
            size_t max_col = 0;
            for (int k = 0; k < kCols; ++k) max_col = std::max(max_col, (size_t)col[k]);
            std::vector<std::string> full;
            auto push = [&](std::string v) {
                full.push_back(std::move(v));
                if (full.size() == n_cols) {
                    std::vector<std::string> r(kCols);
                    for (int k = 0; k < kCols; ++k) r[k] = full[(size_t)col[k]];
                    slow_rows.push_back(std::move(r));
                    full.clear();
                }
            };
            for (const Event* p : atom_site.parts) {
                if (p->kind == EV_TEXT) {
                    std::string v = text_value(x, *p);
                    size_t s = v.size() > 0 && v[0] == ';' ? 1 : 0;
                    size_t e = v.size();
                    while (e > s && v[e - 1] == ';') --e;
                    push(v.substr(s, e - s));
                    continue;
                }
                for_lines(x, p->b, p->e, [&](size_t lb, size_t le) {
                    const u8* b = x.t + lb;
                    const u8* e = x.t + le;
                    trim(b, e);
                    if (b == e || *b == '#') return true;
                    tokenize(b, e, [&](const u8* tb, const u8* te) { push(str(tb, te)); });
                    return true;
                });
            }
            if (!full.empty()) {
                if (max_col >= full.size()) parse_error("incomplete last row in the _atom_site loop");
                std::vector<std::string> r(kCols);
                for (int k = 0; k < kCols; ++k) r[k] = full[(size_t)col[k]];
                slow_rows.push_back(std::move(r));
            }
            n_rows = slow_rows.size();
        }
    } else {
        // no loop: the items _atom_site.<column> form one row
        std::vector<std::string> r(kCols);
        for (int k = 0; k < kCols; ++k) {
            auto it = items.find(std::string("_atom_site.") + kAtomSiteCols[k]);
            if (it == items.end()) parse_error("missing CIF loop key _atom_site.%s", kAtomSiteCols[k]);
            r[k] = it->second;
        }
        slow_rows.push_back(std::move(r));
        n_rows = 1;
    }

    // row visitor: f(row, tokens[16]) over a range of work units
    std::vector<const Event*> runs;
    std::vector<size_t> run_row0;
    if (fast) {
        size_t r0 = 0;
        for (const Event* p : atom_site.parts) {
            runs.push_back(p);
            run_row0.push_back(r0);
            r0 += p->lines;
        }
    }
    const size_t n_units = fast ? runs.size() : (n_rows + 4095) / 4096;
    auto visit = [&](size_t unit, auto&& f) {
        Tok t[kCols];
        if (fast) {
            const Event* p = runs[unit];
            size_t r = run_row0[unit];
            std::vector<Tok> all(n_cols);
            for_lines(x, p->b, p->e, [&](size_t lb, size_t le) {
                const u8* b = x.t + lb;
                const u8* e = x.t + le;
                trim(b, e);
                if (b == e || *b == '#') return true;
                size_t k = 0;
                tokenize(b, e, [&](const u8* tb, const u8* te) { all[k++] = Tok{tb, te}; });
                for (int c = 0; c < kCols; ++c) t[c] = all[(size_t)col[c]];
                f(r++, t);
                return true;
            });
        } else {
            const size_t r1 = std::min(n_rows, (unit + 1) * 4096);
            for (size_t r = unit * 4096; r < r1; ++r) {
                for (int c = 0; c < kCols; ++c) {
                    const u8* b = reinterpret_cast<const u8*>(slow_rows[r][c].data());
                    t[c] = Tok{b, b + slow_rows[r][c].size()};
                }
                f(r, t);
            }
        }
    };

    // 5. model numbers, then model indices in order of first appearance
    Rows rows;
    rows.v.resize(n_rows);
    parallel_for(n_units, nt, [&](size_t u) {
        visit(u, [&](size_t r, const Tok* t) {
            uint64_t m;
            if (!parse_usize(t[13].b, t[13].e, &m))
                parse_error("invalid model number '%s'", str(t[13].b, t[13].e).c_str());
            rows.v[r] = m;
        });
    });
    {
        std::unordered_map<uint64_t, uint32_t> ids;
        uint64_t last_num = 0;
        uint32_t last_idx = 0;
        bool any = false;
        rows.first_phase = n_rows;
        for (size_t r = 0; r < n_rows; ++r) {
            const uint64_t m = rows.v[r];
            uint32_t idx;
            if (any && m == last_num) {
                idx = last_idx;
            } else {
                auto it = ids.find(m);
                if (it == ids.end()) {
                    idx = (uint32_t)ids.size();
                    ids.emplace(m, idx);
                    rows.count.push_back(0);
                    if (idx == 1) rows.first_phase = r;
                } else {
                    idx = it->second;
                }
                last_num = m;
                last_idx = idx;
                any = true;
            }
            const uint64_t pos = rows.count[idx]++;
            rows.v[r] = ((uint64_t)idx << 32) | pos;
        }
    }
    // 6. atoms of the first model
    Parsed ps;
    ps.atoms.resize(rows.first_phase);
    ps.n_models = rows.count.size();
    parallel_for(n_units, nt, [&](size_t u) {
        visit(u, [&](size_t r, const Tok* t) {
            if (r < rows.first_phase) ps.atoms[r] = cif_atom(t);
        });
    });
    std::vector<int32_t> sel;
    finish_structure(ps, out, selection, chain, sel);
    out.n_models = ps.n_models;
    const size_t ns = out.atoms.size();
    out.xyz.assign(out.n_models * ns * 3, 0.0);
    // 7. coordinates (every row is parsed, as the reference does)
    parallel_for(n_units, nt, [&](size_t u) {
        visit(u, [&](size_t r, const Tok* t) {
            double v[3];
            for (int c = 0; c < 3; ++c)
                if (!parse_f64(t[7 + c].b, t[7 + c].e, &v[c]))
                    parse_error("invalid coordinate '%s'", str(t[7 + c].b, t[7 + c].e).c_str());
            const uint64_t m = rows.v[r] >> 32, pos = rows.v[r] & 0xffffffffu;
            if (pos < sel.size() && sel[pos] >= 0) {
                double* o = &out.xyz[(m * ns + (size_t)sel[pos]) * 3];
                o[0] = v[0];
                o[1] = v[1];
                o[2] = v[2];
            }
        });
    });
    check_models(rows.count, ps.atoms.size());
}

// ------------------------------------------------------------ PDB

// &line[i..j]: byte offsets that must lie on character boundaries and inside
// the line (the reference panics otherwise).
inline Tok field(const u8* b, const u8* e, size_t i, size_t j)
{
    const size_t n = (size_t)(e - b);
    if (j > n || i > j) parse_error("PDB line too short: %s", str(b, e).c_str());
    if ((i < n && (b[i] & 0xC0) == 0x80) || (j < n && (b[j] & 0xC0) == 0x80))
        parse_error("PDB field splits a character: %s", str(b, e).c_str());
    return Tok{b + i, b + j};
}

inline Tok trimmed(Tok t)
{
    trim(t.b, t.e);
    return t;
}

// Vec3::from_pdb_line (vec3.rs:241-246).
void pdb_pos(const u8* b, const u8* e, double* v)
{
    for (int c = 0; c < 3; ++c) {
        const Tok t = trimmed(field(b, e, 30 + 8 * (size_t)c, 38 + 8 * (size_t)c));
        if (!parse_f64(t.b, t.e, &v[c])) parse_error("invalid coordinate '%s'", str(t.b, t.e).c_str());
    }
}

// PdbAtom::from_atom_line (pdb_atom.rs:87-116).
AtomRecord pdb_atom(const u8* b, const u8* e)
{
    AtomRecord a;
    Tok t = trimmed(field(b, e, 6, 11));
    if (!parse_i32(t.b, t.e, &a.serial)) parse_error("invalid atom serial '%s'", str(t.b, t.e).c_str());
    t = field(b, e, 12, 16);
    a.name = str(t.b, t.e);
    t = field(b, e, 16, 17);
    a.alt_loc = str(t.b, t.e);
    t = field(b, e, 17, 20);
    a.res_name = str(t.b, t.e);
    t = field(b, e, 21, 22);
    a.chain_id = str(t.b, t.e);
    t = trimmed(field(b, e, 22, 26));
    if (!parse_i32(t.b, t.e, &a.res_seq)) parse_error("invalid residue number '%s'", str(t.b, t.e).c_str());
    t = field(b, e, 26, 27);
    a.i_code = str(t.b, t.e);
    double v[3];
    pdb_pos(b, e, v);
    t = trimmed(field(b, e, 54, 60));
    if (!parse_f64(t.b, t.e, &a.occupancy)) parse_error("invalid occupancy '%s'", str(t.b, t.e).c_str());
    t = trimmed(field(b, e, 60, 66));
    if (!parse_f64(t.b, t.e, &a.temp_factor)) parse_error("invalid B factor '%s'", str(t.b, t.e).c_str());
    if (e - b >= 78) {
        t = trimmed(field(b, e, 77, (size_t)(e - b)));
        a.element = str(t.b, t.e);
        a.has_element = true;
    }
    a.hetero = *b == 'H';
    return a;
}

enum PdbKind : uint8_t { PDB_ATOMS, PDB_MODEL, PDB_ENDMDL };

struct PdbEvent {
    PdbKind kind;
    size_t count = 0;  // PDB_ATOMS: ATOM/HETATM lines in the run
};

// record = line[0..6].trim() for lines of at least 6 bytes (load_pdb.rs:64-67)
inline int pdb_record(const u8* b, const u8* e)
{
    if (e - b < 6) return -1;
    const Tok t = trimmed(field(b, e, 0, 6));
    const size_t n = (size_t)(t.e - t.b);
    if ((n == 4 && std::memcmp(t.b, "ATOM", 4) == 0) || (n == 6 && std::memcmp(t.b, "HETATM", 6) == 0))
        return PDB_ATOMS;
    if (n == 5 && std::memcmp(t.b, "MODEL", 5) == 0) return PDB_MODEL;
    if (n == 6 && std::memcmp(t.b, "ENDMDL", 6) == 0) return PDB_ENDMDL;
    return -1;
}

void pdb_parse(const Text& x, int nt, int selection, const std::string* chain, bs_ensemble& out)
{
    if (first_invalid_line(x) != x.n) parse_error("stream did not contain valid UTF-8");
    const std::vector<Chunk> chunks = make_chunks(x, (size_t)nt * 4);
    const size_t nc = chunks.size();
    std::vector<std::vector<PdbEvent>> events(nc);
    parallel_for(nc, nt, [&](size_t c) {
        auto& ev = events[c];
        for_lines(x, chunks[c].b, chunks[c].e, [&](size_t lb, size_t le) {
            const int k = pdb_record(x.t + lb, x.t + le);
            if (k == PDB_ATOMS) {
                if (ev.empty() || ev.back().kind != PDB_ATOMS) ev.push_back(PdbEvent{PDB_ATOMS});
                ev.back().count += 1;
            } else if (k >= 0) {
                ev.push_back(PdbEvent{(PdbKind)k});
            }
            return true;
        });
    });
    // model bookkeeping (load_pdb.rs:82-107): ENDMDL advances the model id,
    // MODEL opens the next model's list when it is due
    struct RunInfo {
        size_t model, pos;
    };
    std::vector<std::vector<RunInfo>> run_info(nc);
    std::vector<size_t> count(1, 0);
    size_t model_id = 0;
    for (size_t c = 0; c < nc; ++c)
        for (const PdbEvent& e : events[c]) {
            if (e.kind == PDB_ENDMDL) {
                ++model_id;
            } else if (e.kind == PDB_MODEL) {
                if (count.size() == model_id) count.push_back(0);
            } else {
                if (model_id >= count.size()) parse_error("coordinates after ENDMDL without a MODEL record");
                run_info[c].push_back({model_id, count[model_id]});
                count[model_id] += e.count;
            }
        }
    Parsed ps;
    ps.n_models = count.size();
    ps.atoms.resize(count[0]);
    // atoms of the first model, then every line's coordinates
    auto visit = [&](size_t c, auto&& f) {
        size_t ri = 0, in_run = 0;
        bool in_atoms = false;
        for_lines(x, chunks[c].b, chunks[c].e, [&](size_t lb, size_t le) {
            const int k = pdb_record(x.t + lb, x.t + le);
            if (k != PDB_ATOMS) {
                if (k >= 0 && in_atoms) {
                    in_atoms = false;
                    ++ri;
                }
                return true;
            }
            if (!in_atoms) {
                in_atoms = true;
                in_run = 0;
            }
            const RunInfo& r = run_info[c][ri];
            f(r.model, r.pos + in_run, x.t + lb, x.t + le);
            ++in_run;
            return true;
        });
    };
    parallel_for(nc, nt, [&](size_t c) {
        visit(c, [&](size_t m, size_t pos, const u8* b, const u8* e) {
            if (m == 0) ps.atoms[pos] = pdb_atom(b, e);
        });
    });
    std::vector<int32_t> sel;
    finish_structure(ps, out, selection, chain, sel);
    out.n_models = ps.n_models;
    const size_t ns = out.atoms.size();
    out.xyz.assign(out.n_models * ns * 3, 0.0);
    parallel_for(nc, nt, [&](size_t c) {
        visit(c, [&](size_t m, size_t pos, const u8* b, const u8* e) {
            double v[3];
            pdb_pos(b, e, v);
            if (pos < sel.size() && sel[pos] >= 0) {
                double* o = &out.xyz[(m * ns + (size_t)sel[pos]) * 3];
                o[0] = v[0];
                o[1] = v[1];
                o[2] = v[2];
            }
        });
    });
    check_models(count, ps.atoms.size());
}

// ------------------------------------------------------------ entry

bs_ensemble* parse_text(const Text& x0, int format, int selection, const char* chain_id, int n_threads)
{
    BS_REQUIRE(format >= BS_FORMAT_AUTO && format <= BS_FORMAT_PDB, "unknown format %d", format);
    BS_REQUIRE(selection >= BS_SELECT_ALL && selection <= BS_SELECT_NOT_HYDROGEN, "unknown selection %d",
               selection);
    Text x = x0;
    const int fmt = format == BS_FORMAT_AUTO ? detect_format(x) : format;
    std::string chain;
    if (chain_id) chain = chain_id;
    auto ens = std::make_unique<bs_ensemble>();
    ens->format = fmt;
    const int nt = thread_count(n_threads, x.n);
    if (fmt == BS_FORMAT_CIF) {
        x.n = first_invalid_line(x);  // the reader stops at a line that is not UTF-8
        cif_parse(x, nt, selection, chain_id ? &chain : nullptr, *ens);
    } else {
        pdb_parse(x, nt, selection, chain_id ? &chain : nullptr, *ens);
    }
    return ens.release();
}

// Whole-file read: mapped when plain, inflated when the name ends in .gz
// (bioshell-core/src/io/utils.rs:312-330).
struct FileText {
    std::vector<u8> owned;
    void* map = nullptr;
    size_t map_len = 0;
    Text text{nullptr, 0};
    ~FileText()
    {
        if (map) munmap(map, map_len);
    }
};

bool gz_name(const std::string& path)
{
    const size_t slash = path.find_last_of('/');
    const std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    return dot != std::string::npos && dot > 0 && name.substr(dot + 1) == "gz";
}

void read_file(const std::string& path, FileText& f)
{
    if (path.empty()) fail(BS_ERR_INVALID_ARGUMENT, "empty file name");
    if (gz_name(path)) {
        gzFile g = gzopen(path.c_str(), "rb");
        if (!g) fail(BS_ERR_IO, "cannot open %s", path.c_str());
        gzbuffer(g, 1 << 20);
        std::vector<u8>& buf = f.owned;
        size_t n = 0;
        for (;;) {
            if (buf.size() - n < (1u << 20)) buf.resize(std::max<size_t>(buf.size() * 2, n + (4u << 20)));
            const int r = gzread(g, buf.data() + n, (unsigned)std::min<size_t>(buf.size() - n, 1u << 30));
            if (r < 0) {
                int errnum = 0;
                const std::string msg = gzerror(g, &errnum);
                gzclose(g);
                fail(BS_ERR_IO, "cannot inflate %s: %s", path.c_str(), msg.c_str());
            }
            if (r == 0) break;
            n += (size_t)r;
        }
        gzclose(g);
        buf.resize(n);
        f.text = Text{buf.data(), n};
        return;
    }
    const int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) fail(BS_ERR_IO, "cannot open %s: %s", path.c_str(), std::strerror(errno));
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        fail(BS_ERR_IO, "cannot stat %s", path.c_str());
    }
    const size_t n = (size_t)st.st_size;
    if (n > 0) {
        void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            close(fd);
            fail(BS_ERR_IO, "cannot map %s", path.c_str());
        }
        madvise(m, n, MADV_SEQUENTIAL);
        f.map = m;
        f.map_len = n;
        f.text = Text{static_cast<const u8*>(m), n};
    }
    close(fd);
}

void copy_str(char* dst, size_t cap, const std::string& s)
{
    size_t n = std::min(s.size(), cap - 1);
    while (n > 0 && n < s.size() && (static_cast<unsigned char>(s[n]) & 0xC0) == 0x80) --n;
    std::memcpy(dst, s.data(), n);
    dst[n] = 0;
}

// Rust's {:>w} / {:^w} / {:w} count characters, not bytes.
size_t chars_of(const std::string& s)
{
    size_t n = 0;
    for (unsigned char c : s) n += (c & 0xC0) != 0x80;
    return n;
}

void pad_right_align(std::string& out, const std::string& s, size_t w)
{
    const size_t n = chars_of(s);
    if (n < w) out.append(w - n, ' ');
    out += s;
}

void pad_left_align(std::string& out, const std::string& s, size_t w)
{
    const size_t n = chars_of(s);
    out += s;
    if (n < w) out.append(w - n, ' ');
}

void pad_centre(std::string& out, const std::string& s, size_t w)
{
    const size_t n = chars_of(s);
    const size_t pad = n < w ? w - n : 0;
    out.append(pad / 2, ' ');
    out += s;
    out.append(pad - pad / 2, ' ');
}

// {:>w.p} of an f64 (Rust prints NaN as "NaN" and infinities as "inf").
void fixed(std::string& out, double v, int w, int p)
{
    char buf[64];
    if (std::isnan(v))
        std::snprintf(buf, sizeof buf, "%*s", w, "NaN");
    else if (std::isinf(v))
        std::snprintf(buf, sizeof buf, "%*s", w, v < 0 ? "-inf" : "inf");
    else
        std::snprintf(buf, sizeof buf, "%*.*f", w, p, v);
    out += buf;
}

}  // namespace

std::string pdb_atom_line(const AtomRecord& a, const std::string& chain, double x, double y, double z)
{
    std::string s;
    s.reserve(84);
    s += a.hetero ? "HETATM" : "ATOM  ";
    pad_right_align(s, std::to_string(a.serial), 5);
    s += ' ';
    pad_centre(s, a.name, 4);
    pad_right_align(s, a.alt_loc, 1);
    pad_right_align(s, a.res_name, 3);
    s += ' ';
    pad_right_align(s, chain, 1);
    pad_right_align(s, std::to_string(a.res_seq), 4);
    pad_left_align(s, a.i_code, 1);
    s += "   ";
    fixed(s, x, 8, 3);
    fixed(s, y, 8, 3);
    fixed(s, z, 8, 3);
    fixed(s, a.occupancy, 6, 2);
    fixed(s, a.temp_factor, 6, 2);
    s += "          ";
    pad_right_align(s, a.has_element ? a.element : std::string("  "), 2);
    s += "  ";
    return s;
}

void append_pdb_model(std::string& out, const std::vector<AtomRecord>& atoms, const double* xyz)
{
    std::vector<std::string> chains;
    for (const AtomRecord& a : atoms) chains.push_back(a.chain_id);
    std::sort(chains.begin(), chains.end());
    chains.erase(std::unique(chains.begin(), chains.end()), chains.end());
    bool rename = false;
    for (const std::string& c : chains) rename = rename || c.size() > 1;
    static const char kCodes[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789";
    if (rename && chains.size() > 36) fail(BS_ERR_INVALID_ARGUMENT, "more than 36 chains to rename");
    for (size_t i = 0; i < chains.size(); ++i) {
        const std::string code = rename ? std::string(1, kCodes[i]) : chains[i];
        for (size_t k = 0; k < atoms.size(); ++k)
            if (atoms[k].chain_id == chains[i]) {
                out += pdb_atom_line(atoms[k], code, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]);
                out += '\n';
            }
    }
}

}  // namespace bs

using namespace bs;

extern "C" {

bs_status bs_ensemble_parse(const char* text, size_t len, int format, int selection, const char* chain_id,
                            int n_threads, bs_ensemble** out)
{
    return guard([&] {
        BS_REQUIRE(out, "null output pointer");
        *out = nullptr;
        BS_REQUIRE(text || len == 0, "null text");
        const Text x{reinterpret_cast<const unsigned char*>(text), len};
        *out = parse_text(x, format, selection, chain_id, n_threads);
    });
}

bs_status bs_ensemble_load(const char* path, int format, int selection, const char* chain_id, int n_threads,
                           bs_ensemble** out)
{
    return guard([&] {
        BS_REQUIRE(out, "null output pointer");
        *out = nullptr;
        BS_REQUIRE(path, "null path");
        FileText f;
        read_file(path, f);
        *out = parse_text(f.text, format, selection, chain_id, n_threads);
    });
}

void bs_ensemble_destroy(bs_ensemble* e) { delete e; }

bs_status bs_ensemble_get_info(const bs_ensemble* e, bs_ensemble_info* info)
{
    return guard([&] {
        BS_REQUIRE(e && info, "null pointer argument");
        info->n_models = e->n_models;
        info->n_atoms = e->atoms.size();
        info->n_structure_atoms = e->n_structure_atoms;
        info->n_chains = e->n_chains;
        info->n_residues = e->n_residues;
        info->format = e->format;
    });
}

bs_status bs_ensemble_copy_coords(const bs_ensemble* e, double* xyz)
{
    return guard([&] {
        BS_REQUIRE(e, "null ensemble");
        BS_REQUIRE(xyz || e->xyz.empty(), "null output");
        if (!e->xyz.empty()) std::memcpy(xyz, e->xyz.data(), e->xyz.size() * sizeof(double));
    });
}

bs_status bs_ensemble_get_atom(const bs_ensemble* e, size_t k, bs_atom_record* out)
{
    return guard([&] {
        BS_REQUIRE(e && out, "null pointer argument");
        BS_REQUIRE(k < e->atoms.size(), "atom index %zu out of range (%zu atoms)", k, e->atoms.size());
        const AtomRecord& a = e->atoms[k];
        out->serial = a.serial;
        out->res_seq = a.res_seq;
        out->occupancy = a.occupancy;
        out->temp_factor = a.temp_factor;
        copy_str(out->name, sizeof out->name, a.name);
        copy_str(out->alt_loc, sizeof out->alt_loc, a.alt_loc);
        copy_str(out->res_name, sizeof out->res_name, a.res_name);
        copy_str(out->chain_id, sizeof out->chain_id, a.chain_id);
        copy_str(out->i_code, sizeof out->i_code, a.i_code);
        copy_str(out->element, sizeof out->element, a.element);
        out->has_element = a.has_element ? 1 : 0;
        out->hetero = a.hetero ? 1 : 0;
    });
}

bs_status bs_ensemble_write_pdb(const bs_ensemble* e, const double* xyz, size_t n_models, const char* path)
{
    return guard([&] {
        BS_REQUIRE(e && path, "null pointer argument");
        const double* src = xyz ? xyz : e->xyz.data();
        const size_t nm = xyz ? n_models : e->n_models;
        BS_REQUIRE(nm >= 1, "nothing to write");
        const size_t na = e->atoms.size();
        std::string out;
        out.reserve(nm * (na * 81 + 32));
        for (size_t m = 0; m < nm; ++m) {
            if (nm > 1) {
                char buf[32];
                std::snprintf(buf, sizeof buf, "MODEL     %4zu\n", m + 1);
                out += buf;
            }
            append_pdb_model(out, e->atoms, src + m * na * 3);
            if (nm > 1) out += "ENDMDL\n";
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) fail(BS_ERR_IO, "cannot open %s for writing", path);
        const size_t w = std::fwrite(out.data(), 1, out.size(), f);
        const int rc = std::fclose(f);
        if (w != out.size() || rc != 0) fail(BS_ERR_IO, "cannot write %s", path);
    });
}

bs_status bs_ensemble_upload(bs_context* ctx, const bs_ensemble* e, double* d_xyz)
{
    return guard([&] {
        BS_REQUIRE(ctx && e, "null pointer argument");
        if (e->xyz.empty()) return;
        BS_REQUIRE(d_xyz, "null device pointer");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        BS_HIP(hipMemcpyAsync(d_xyz, e->xyz.data(), e->xyz.size() * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
