// C++ facade over include/cqbem.h with the reference's collection interface
// (src/elastodynamics_cqbem_matrix_collection.cpp:13-278, src/physics_base_matrix_collection.cpp:32-57,
// src/elastics_bem_matrix_collection.cpp:16-41).
#include "elastodynamics_cqbem_matrix_collection.hpp"

#include "cqbem_rules.hpp"
#include "host_parallel.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <functional>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace cqbem {

// ----------------------------------------------------------------- JSON --
namespace {

struct Parser {
    const std::string &s;
    size_t p = 0;
    explicit Parser(const std::string &t) : s(t) {}
    [[noreturn]] void err(const char *what) { throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(p)); }
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\n' || s[p] == '\t' || s[p] == '\r')) p++;
    }
    bool lit(const char *w) {
        const size_t n = std::char_traits<char>::length(w);
        if (s.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    unsigned hex4() {
        if (p + 4 > s.size()) err("bad \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; k++) {
            const char c = s[p++];
            v = v * 16 + (c >= '0' && c <= '9' ? c - '0' : c >= 'a' && c <= 'f' ? c - 'a' + 10 : c >= 'A' && c <= 'F' ? c - 'A' + 10 : (err("bad \\u escape"), 0));
        }
        return v;
    }
    static void utf8(std::string &out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | cp >> 6), out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | cp >> 12), out += (char)(0x80 | (cp >> 6 & 0x3F)), out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | cp >> 18), out += (char)(0x80 | (cp >> 12 & 0x3F));
            out += (char)(0x80 | (cp >> 6 & 0x3F)), out += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string str() {
        if (s[p] != '"') err("expected string");
        p++;
        std::string out;
        while (p < s.size() && s[p] != '"') {
            if (s[p] != '\\') {
                out += s[p++];
                continue;
            }
            if (++p >= s.size()) err("bad escape");
            const char c = s[p++];
            switch (c) {
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case '"': case '\\': case '/': out += c; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
                    if (!lit("\\u")) err("unpaired surrogate");
                    const unsigned lo = hex4();
                    if (lo < 0xDC00 || lo >= 0xE000) err("unpaired surrogate");
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(out, cp);
                break;
            }
            default: err("bad escape");
            }
        }
        if (p >= s.size()) err("unterminated string");
        p++;
        return out;
    }
    Json value() {
        ws();
        if (p >= s.size()) err("unexpected end");
        Json j;
        const char c = s[p];
        if (c == '{') {
            j.kind = Json::Object;
            p++;
            ws();
            if (s[p] == '}') return p++, j;
            for (;;) {
                ws();
                std::string k = str();
                ws();
                if (s[p++] != ':') err("expected ':'");
                j.obj[k] = value();
                ws();
                if (s[p] == ',') {
                    p++;
                    continue;
                }
                if (s[p] == '}') return p++, j;
                err("expected ',' or '}'");
            }
        }
        if (c == '[') {
            j.kind = Json::Array;
            p++;
            ws();
            if (s[p] == ']') return p++, j;
            for (;;) {
                j.arr.push_back(value());
                ws();
                if (s[p] == ',') {
                    p++;
                    continue;
                }
                if (s[p] == ']') return p++, j;
                err("expected ',' or ']'");
            }
        }
        if (c == '"') {
            j.kind = Json::String;
            j.str = str();
            return j;
        }
        if (lit("true")) return j.kind = Json::Bool, j.b = true, j;
        if (lit("false")) return j.kind = Json::Bool, j.b = false, j;
        if (lit("null")) return j;
        char *end = nullptr;
        const char *start = s.c_str() + p;
        j.num = std::strtod(start, &end);
        if (end == start) err("bad value");
        const std::string tok(start, static_cast<const char *>(end));
        if (tok.find_first_of(".eE") == std::string::npos) {  // an integer literal, if it fits
            errno = 0;
            const long long v = std::strtoll(tok.c_str(), nullptr, 10);
            if (errno == 0) j.integer = true, j.inum = v;
        }
        p = end - s.c_str();
        j.kind = Json::Number;
        return j;
    }
};

}  // namespace

Json Json::parse(const std::string &text) {
    Parser ps(text);
    Json j = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.err("trailing characters");
    return j;
}

const Json &Json::operator[](const std::string &k) const {
    static const Json null_value;
    if (kind != Object) throw std::runtime_error("json: not an object when reading '" + k + "'");
    auto it = obj.find(k);
    if (it == obj.end()) throw std::runtime_error("json: missing key '" + k + "'");
    return it->second;
    (void)null_value;
}
double Json::number() const {
    if (kind != Number) throw std::runtime_error("json: expected a number");
    return num;
}
const std::string &Json::string() const {
    if (kind != String) throw std::runtime_error("json: expected a string");
    return str;
}
bool Json::boolean() const {
    if (kind != Bool) throw std::runtime_error("json: expected a boolean");
    return b;
}

namespace {

// Shortest-digit generation of the reference's JSON library for doubles: Grisu2 (F. Loitsch, "Printing
// floating-point numbers quickly and accurately with integers", PLDI 2010) with the scaled boundaries narrowed
// by one unit, so that the digits read back to the same double; most of the time they are the shortest
// such digits, and where Grisu2 emits a longer string the cache key must contain that same string.
//
// The digit generation, the cached-power table layout and the rounding loop below follow the Grisu2
// implementation of nlohmann/json 3.9.1 (the reference's vendored include/json.hpp:14600-15300), because the
// cache key must be byte-identical to the one that library prints.  That code carries this notice:
//   JSON for Modern C++, https://github.com/nlohmann/json
//   Licensed under the MIT License <http://opensource.org/licenses/MIT>.  SPDX-License-Identifier: MIT
//   Copyright (c) 2013-2019 Niels Lohmann <http://nlohmann.me>.
//   Permission is hereby granted, free of charge, to any person obtaining a copy of this software and
//   associated documentation files (the "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish, distribute, sublicense, and/or sell
//   copies of the Software, and to permit persons to whom the Software is furnished to do so, subject to the
//   following conditions: The above copyright notice and this permission notice shall be included in all
//   copies or substantial portions of the Software.
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR IMPLIED, INCLUDING BUT NOT
//   LIMITED TO THE WARRANTIES OF MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO
//   EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER LIABILITY, WHETHER IN
//   AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
struct DiyFp {
    uint64_t f;
    int e;
};

DiyFp diy_mul(DiyFp x, DiyFp y) {  // upper 64 bits of the 128-bit product, rounded half up
    const unsigned __int128 p = (unsigned __int128)x.f * y.f;
    const uint64_t hi = (uint64_t)(p >> 64), lo = (uint64_t)p;
    return {hi + (lo >> 63), x.e + y.e + 64};
}

DiyFp diy_normalize(DiyFp x) {
    while ((x.f >> 63) == 0) x.f <<= 1, x.e--;
    return x;
}

// 10^k for k = -300, -292, ..., 324: the significand rounded to 64 bits and its binary exponent (computed
// from the exact powers of ten).
struct CachedPower {
    uint64_t f;
    int e, k;
};
constexpr CachedPower kCachedPowers[] = {
    {0xAB70FE17C79AC6CA, -1060, -300}, {0xFF77B1FCBEBCDC4F, -1034, -292},
    {0xBE5691EF416BD60C, -1007, -284}, {0x8DD01FAD907FFC3C, -980, -276},
    {0xD3515C2831559A83, -954, -268}, {0x9D71AC8FADA6C9B5, -927, -260},
    {0xEA9C227723EE8BCB, -901, -252}, {0xAECC49914078536D, -874, -244},
    {0x823C12795DB6CE57, -847, -236}, {0xC21094364DFB5637, -821, -228},
    {0x9096EA6F3848984F, -794, -220}, {0xD77485CB25823AC7, -768, -212},
    {0xA086CFCD97BF97F4, -741, -204}, {0xEF340A98172AACE5, -715, -196},
    {0xB23867FB2A35B28E, -688, -188}, {0x84C8D4DFD2C63F3B, -661, -180},
    {0xC5DD44271AD3CDBA, -635, -172}, {0x936B9FCEBB25C996, -608, -164},
    {0xDBAC6C247D62A584, -582, -156}, {0xA3AB66580D5FDAF6, -555, -148},
    {0xF3E2F893DEC3F126, -529, -140}, {0xB5B5ADA8AAFF80B8, -502, -132},
    {0x87625F056C7C4A8B, -475, -124}, {0xC9BCFF6034C13053, -449, -116},
    {0x964E858C91BA2655, -422, -108}, {0xDFF9772470297EBD, -396, -100},
    {0xA6DFBD9FB8E5B88F, -369, -92}, {0xF8A95FCF88747D94, -343, -84},
    {0xB94470938FA89BCF, -316, -76}, {0x8A08F0F8BF0F156B, -289, -68},
    {0xCDB02555653131B6, -263, -60}, {0x993FE2C6D07B7FAC, -236, -52},
    {0xE45C10C42A2B3B06, -210, -44}, {0xAA242499697392D3, -183, -36},
    {0xFD87B5F28300CA0E, -157, -28}, {0xBCE5086492111AEB, -130, -20},
    {0x8CBCCC096F5088CC, -103, -12}, {0xD1B71758E219652C, -77, -4},
    {0x9C40000000000000, -50, 4}, {0xE8D4A51000000000, -24, 12},
    {0xAD78EBC5AC620000, 3, 20}, {0x813F3978F8940984, 30, 28},
    {0xC097CE7BC90715B3, 56, 36}, {0x8F7E32CE7BEA5C70, 83, 44},
    {0xD5D238A4ABE98068, 109, 52}, {0x9F4F2726179A2245, 136, 60},
    {0xED63A231D4C4FB27, 162, 68}, {0xB0DE65388CC8ADA8, 189, 76},
    {0x83C7088E1AAB65DB, 216, 84}, {0xC45D1DF942711D9A, 242, 92},
    {0x924D692CA61BE758, 269, 100}, {0xDA01EE641A708DEA, 295, 108},
    {0xA26DA3999AEF774A, 322, 116}, {0xF209787BB47D6B85, 348, 124},
    {0xB454E4A179DD1877, 375, 132}, {0x865B86925B9BC5C2, 402, 140},
    {0xC83553C5C8965D3D, 428, 148}, {0x952AB45CFA97A0B3, 455, 156},
    {0xDE469FBD99A05FE3, 481, 164}, {0xA59BC234DB398C25, 508, 172},
    {0xF6C69A72A3989F5C, 534, 180}, {0xB7DCBF5354E9BECE, 561, 188},
    {0x88FCF317F22241E2, 588, 196}, {0xCC20CE9BD35C78A5, 614, 204},
    {0x98165AF37B2153DF, 641, 212}, {0xE2A0B5DC971F303A, 667, 220},
    {0xA8D9D1535CE3B396, 694, 228}, {0xFB9B7CD9A4A7443C, 720, 236},
    {0xBB764C4CA7A44410, 747, 244}, {0x8BAB8EEFB6409C1A, 774, 252},
    {0xD01FEF10A657842C, 800, 260}, {0x9B10A4E5E9913129, 827, 268},
    {0xE7109BFBA19C0C9D, 853, 276}, {0xAC2820D9623BF429, 880, 284},
    {0x80444B5E7AA7CF85, 907, 292}, {0xBF21E44003ACDD2D, 933, 300},
    {0x8E679C2F5E44FF8F, 960, 308}, {0xD433179D9C8CB841, 986, 316},
    {0x9E19DB92B4E31BA9, 1013, 324},
};

// Digits of v (finite, > 0) into buf; v = digits x 10^dec_exp.
void grisu2(double v, char *buf, int &len, int &dec_exp) {
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    const uint64_t F = bits & ((1ULL << 52) - 1);
    const int E = (int)(bits >> 52);
    const DiyFp x = E == 0 ? DiyFp{F, 1 - 1075} : DiyFp{F | (1ULL << 52), E - 1075};
    // the half-way points to the neighbouring doubles; the lower one is closer at a power of two
    const bool closer = F == 0 && E > 1;
    const DiyFp m_plus = diy_normalize({2 * x.f + 1, x.e - 1});
    DiyFp m_minus = closer ? DiyFp{4 * x.f - 1, x.e - 2} : DiyFp{2 * x.f - 1, x.e - 1};
    m_minus = {m_minus.f << (m_minus.e - m_plus.e), m_plus.e};
    const DiyFp w = diy_normalize(x);
    // cached power bringing the product's binary exponent into [-60, -32]: the smallest tabulated decimal
    // exponent >= ceil((-60 - e - 1) log10 2), log10 2 taken as 78913 / 2^18
    const int fexp = -60 - m_plus.e - 1;
    const int kmin = (fexp * 78913) / (1 << 18) + (fexp > 0 ? 1 : 0);
    const CachedPower c = kCachedPowers[(300 + kmin + 7) / 8];
    const DiyFp cp{c.f, c.e};
    const DiyFp W = diy_mul(w, cp), Wm = diy_mul(m_minus, cp), Wp = diy_mul(m_plus, cp);
    const DiyFp lo{Wm.f + 1, Wm.e}, hi{Wp.f - 1, Wp.e};  // narrowed by one unit: the scaling is inexact
    dec_exp = -c.k;
    // digits of hi, stopped once hi - digits-so-far is within delta = hi - lo
    uint64_t delta = hi.f - lo.f, dist = hi.f - W.f;
    const int sh = -hi.e;
    const uint64_t one = 1ULL << sh;
    uint32_t p1 = (uint32_t)(hi.f >> sh);
    uint64_t p2 = hi.f & (one - 1);
    auto round_last = [&](uint64_t rest, uint64_t ten_k) {  // move the last digit towards W while it stays in range
        while (rest < dist && delta - rest >= ten_k && (rest + ten_k < dist || dist - rest > rest + ten_k - dist)) {
            buf[len - 1]--;
            rest += ten_k;
        }
    };
    len = 0;
    uint32_t pow10 = 1;
    int n = 1;
    while (n < 10 && pow10 * 10ULL <= p1) pow10 *= 10, n++;
    while (n > 0) {
        buf[len++] = (char)('0' + p1 / pow10);
        p1 %= pow10;
        n--;
        const uint64_t rest = ((uint64_t)p1 << sh) + p2;
        if (rest <= delta) {
            dec_exp += n;
            round_last(rest, (uint64_t)pow10 << sh);
            return;
        }
        pow10 /= 10;
    }
    int m = 0;
    for (;;) {
        p2 *= 10;
        buf[len++] = (char)('0' + (p2 >> sh));
        p2 &= one - 1;
        m++;
        delta *= 10;
        dist *= 10;
        if (p2 <= delta) break;
    }
    dec_exp -= m;
    round_last(p2, one);
}

// A finite double as the reference's JSON library prints it: the Grisu2 digits, then fixed notation for
// decimal exponents in (-4, 15] (with ".0" on integral values) and d.ddde+XX otherwise.
std::string format_double(double v) {
    if (!std::isfinite(v)) return "null";
    std::string out = std::signbit(v) ? "-" : "";
    v = std::fabs(v);
    if (v == 0.0) return out + "0.0";
    char buf[32];
    int len = 0, dec = 0;
    grisu2(v, buf, len, dec);
    const std::string digits(buf, buf + len);
    const int k = len, n = len + dec;  // value = 0.digits x 10^n
    constexpr int kMinExp = -4, kMaxExp = 15;
    if (k <= n && n <= kMaxExp) return out + digits + std::string(n - k, '0') + ".0";
    if (0 < n && n <= kMaxExp) return out + digits.substr(0, n) + "." + digits.substr(n);
    if (kMinExp < n && n <= 0) return out + "0." + std::string(-n, '0') + digits;
    std::string m = k == 1 ? digits : digits.substr(0, 1) + "." + digits.substr(1);
    const int x = n - 1;
    char ex[8];
    std::snprintf(ex, sizeof(ex), "%c%02d", x < 0 ? '-' : '+', x < 0 ? -x : x);
    return out + m + "e" + ex;
}

void dump_string(std::string &o, const std::string &str) {
    o += '"';
    for (unsigned char c : str) {
        switch (c) {
        case '\b': o += "\\b"; break;
        case '\t': o += "\\t"; break;
        case '\n': o += "\\n"; break;
        case '\f': o += "\\f"; break;
        case '\r': o += "\\r"; break;
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        default:
            if (c < 0x20) {
                char u[8];
                std::snprintf(u, sizeof(u), "\\u%04x", c);
                o += u;
            } else {
                o += (char)c;
            }
        }
    }
    o += '"';
}

void dump_value(std::string &o, const Json &j, int indent, int level) {
    const bool pretty = indent >= 0;
    auto nl = [&](int lv) {
        if (pretty) o += '\n' + std::string((size_t)(indent * lv), ' ');
    };
    switch (j.kind) {
    case Json::Null: o += "null"; return;
    case Json::Bool: o += j.b ? "true" : "false"; return;
    case Json::Number: o += j.integer ? std::to_string(j.inum) : format_double(j.num); return;
    case Json::String: dump_string(o, j.str); return;
    case Json::Array:
        if (j.arr.empty()) {
            o += "[]";
            return;
        }
        o += '[';
        for (size_t k = 0; k < j.arr.size(); k++) {
            if (k) o += ',';
            nl(level + 1);
            dump_value(o, j.arr[k], indent, level + 1);
        }
        nl(level);
        o += ']';
        return;
    case Json::Object:
        if (j.obj.empty()) {
            o += "{}";
            return;
        }
        o += '{';
        bool first = true;
        for (const auto &kv : j.obj) {
            if (!first) o += ',';
            first = false;
            nl(level + 1);
            dump_string(o, kv.first);
            o += pretty ? ": " : ":";
            dump_value(o, kv.second, indent, level + 1);
        }
        nl(level);
        o += '}';
        return;
    }
}

}  // namespace

std::string Json::dump(int indent) const {
    std::string o;
    dump_value(o, *this, indent, 0);
    return o;
}

std::string cache_key(const Json &group, Json *stripped) {
    Json copy = group;
    if (copy.contains("simulation")) {
        auto &sim = copy.obj["simulation"].obj;
        sim.erase("compression_ratio");
        sim.erase("update_local_frame");
    }
    std::ostringstream ss;
    ss << std::uppercase << std::hex << std::hash<std::string>()(copy.dump());
    if (stripped) *stripped = std::move(copy);
    return ss.str();
}

bool save_matrix(const Dense &m, const std::string &path) {
    std::ofstream out(path, std::ios::binary);
    if (!out) return false;
    const int64_t rc[2] = {m.rows, m.cols};
    out.write(reinterpret_cast<const char *>(rc), sizeof(rc));
    out.write(reinterpret_cast<const char *>(m.data.data()), (std::streamsize)(m.data.size() * sizeof(double)));
    return (bool)out;
}

bool load_matrix(Dense &m, const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    int64_t rc[2] = {0, 0};
    in.read(reinterpret_cast<char *>(rc), sizeof(rc));
    if (!in || rc[0] < 0 || rc[1] < 0) return false;
    m = Dense(rc[0], rc[1]);
    in.read(reinterpret_cast<char *>(m.data.data()), (std::streamsize)(m.data.size() * sizeof(double)));
    return (bool)in;
}

// ----------------------------------------------------------------- .cpmat --
bool save_cpmat(const DeviceCsr &m, const std::string &path) {
    std::vector<int64_t> rp((size_t)m.rows + 1);
    std::vector<int32_t> col((size_t)m.nnz);
    std::vector<double> val((size_t)m.nnz);
    if (hipMemcpy(rp.data(), m.rowptr, sizeof(int64_t) * rp.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        (m.nnz && (hipMemcpy(col.data(), m.col, sizeof(int32_t) * m.nnz, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(val.data(), m.val, sizeof(double) * m.nnz, hipMemcpyDeviceToHost) != hipSuccess)))
        return false;
    std::ofstream out(path, std::ios::binary);
    if (!out) return false;
    const int32_t N = (int32_t)m.n;
    const int64_t head[5] = {m.rows, m.rows, m.nnz, m.rows, m.rows};  // rows, cols, nonZeros, outerSize, innerSize
    out.write(reinterpret_cast<const char *>(&N), sizeof N);
    out.write(reinterpret_cast<const char *>(head), sizeof head);
    out.write(reinterpret_cast<const char *>(val.data()), sizeof(double) * m.nnz);
    std::vector<int32_t> outer(rp.begin(), rp.end() - 1);
    out.write(reinterpret_cast<const char *>(outer.data()), sizeof(int32_t) * outer.size());
    out.write(reinterpret_cast<const char *>(col.data()), sizeof(int32_t) * m.nnz);
    return (bool)out;
}

bool load_cpmat(DeviceCsr &m, int64_t n, const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    int32_t N = 0;
    int64_t head[5];
    in.read(reinterpret_cast<char *>(&N), sizeof N);
    in.read(reinterpret_cast<char *>(head), sizeof head);
    const int64_t rows = 2 * wavelet_padded(n) - 2, nnz = head[2];
    if (!in || N != n || head[0] != rows || head[1] != rows || head[3] != rows || head[4] != rows || nnz < 0 ||
        nnz > rows * rows)
        return false;
    std::vector<double> val((size_t)nnz);
    std::vector<int32_t> outer((size_t)rows), col((size_t)nnz);
    in.read(reinterpret_cast<char *>(val.data()), sizeof(double) * nnz);
    in.read(reinterpret_cast<char *>(outer.data()), sizeof(int32_t) * rows);
    in.read(reinterpret_cast<char *>(col.data()), sizeof(int32_t) * nnz);
    if (!in) return false;
    std::vector<int64_t> rp((size_t)rows + 1);
    for (int64_t r = 0; r < rows; r++) rp[r] = outer[r];
    rp[rows] = nnz;
    for (int64_t r = 0; r < rows; r++)
        if (rp[r] < 0 || rp[r] > rp[r + 1]) return false;
    for (int32_t c : col)
        if (c < 0 || c >= rows) return false;
    m.release();
    if (hipMalloc(&m.rowptr, sizeof(int64_t) * (rows + 1)) != hipSuccess ||
        hipMalloc(&m.col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
        hipMalloc(&m.val, sizeof(double) * std::max<int64_t>(nnz, 1)) != hipSuccess) {
        m.release();
        return false;
    }
    m.n = n, m.rows = rows, m.nnz = nnz;
    if (hipMemcpy(m.rowptr, rp.data(), sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (nnz && (hipMemcpy(m.col, col.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice) != hipSuccess ||
                 hipMemcpy(m.val, val.data(), sizeof(double) * nnz, hipMemcpyHostToDevice) != hipSuccess))) {
        m.release();
        return false;
    }
    return true;
}

// ------------------------------------------------------------------- OBJ --
bool load_obj(const std::string &path, std::vector<double> &V, std::vector<int32_t> &F, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open mesh file " + path;
        return false;
    }
    V.clear(), F.clear();
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            double x, y, z;
            ls >> x >> y >> z;
            V.insert(V.end(), {x, y, z});
        } else if (tag == "f") {
            std::vector<int> idx;
            std::string tok;
            while (ls >> tok) {
                const int k = std::stoi(tok.substr(0, tok.find('/')));
                idx.push_back(k > 0 ? k - 1 : (int)(V.size() / 3) + k);
            }
            for (size_t k = 1; k + 1 < idx.size(); k++) F.insert(F.end(), {idx[0], idx[k], idx[k + 1]});
        }
    }
    if (V.empty() || F.empty()) {
        err = "empty mesh in " + path;
        return false;
    }
    const int32_t nv = (int32_t)(V.size() / 3);
    for (int32_t x : F)
        if (x < 0 || x >= nv) {
            err = "face index out of range in " + path;
            return false;
        }
    return true;
}

// ------------------------------------------------------------- collection --
namespace {
// Device-to-host copies of the assembled matrices through two pinned staging buffers: the DMA of one chunk runs
// while the host threads copy the previous chunk into its pageable destination (negated for H).  A pageable
// hipMemcpy reaches about 18 GB/s here.
class DownloadStager {
  public:
    static constexpr int64_t kChunk = (int64_t)1 << 23;  // doubles per chunk (64 MB)
    DownloadStager() {
        ok_ = hipStreamCreateWithFlags(&st_, hipStreamNonBlocking) == hipSuccess;
        for (int k = 0; k < 2 && ok_; k++)
            ok_ = hipHostMalloc((void **)&pin_[k], sizeof(double) * kChunk, hipHostMallocDefault) == hipSuccess &&
                  hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming) == hipSuccess;
    }
    ~DownloadStager() {
        if (st_) (void)hipStreamSynchronize(st_);
        for (int k = 0; k < 2; k++) {
            if (pin_[k]) (void)hipHostFree(pin_[k]);
            if (ev_[k]) (void)hipEventDestroy(ev_[k]);
        }
        if (st_) (void)hipStreamDestroy(st_);
    }
    DownloadStager(const DownloadStager &) = delete;
    DownloadStager &operator=(const DownloadStager &) = delete;
    bool ok() const { return ok_; }

    // dst[i] = sign * src[i], i < n (src in device memory); false on a HIP error.
    bool copy(double *dst, const double *src, int64_t n, double sign) {
        const int64_t chunks = (n + kChunk - 1) / kChunk;
        auto issue = [&](int64_t c) {
            const int64_t len = std::min(kChunk, n - c * kChunk);
            return hipMemcpyAsync(pin_[c & 1], src + c * kChunk, sizeof(double) * len, hipMemcpyDeviceToHost, st_) == hipSuccess &&
                   hipEventRecord(ev_[c & 1], st_) == hipSuccess;
        };
        if (chunks > 0 && !issue(0)) return false;
        for (int64_t c = 0; c < chunks; c++) {
            if (c + 1 < chunks && !issue(c + 1)) return false;  // its buffer was consumed by the host at c - 1
            if (hipEventSynchronize(ev_[c & 1]) != hipSuccess) return false;
            const int64_t len = std::min(kChunk, n - c * kChunk);
            const double *s = pin_[c & 1];
            double *d = dst + c * kChunk;
            parallel_for(len, [&](int64_t a, int64_t b) {
                if (sign == 1.0)
                    std::memcpy(d + a, s + a, sizeof(double) * (size_t)(b - a));
                else
                    for (int64_t e = a; e < b; e++) d[e] = -s[e];
            });
        }
        return true;
    }

  private:
    bool ok_ = false;
    hipStream_t st_ = nullptr;
    double *pin_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
};

std::mutex g_ctx_mu;
cqbem_context *g_ctx = nullptr;  // one GPU context per process, shared by all collections

cqbem_context *context(std::string &err) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx && cqbem_create(&g_ctx, 0, 0) != CQBEM_OK) {
        g_ctx = nullptr;
        err = "no HIP device available for the CQBEM precompute";
    }
    return g_ctx;
}
}  // namespace

ElastodynamicsCQBEMMatrixCollection::ElastodynamicsCQBEMMatrixCollection(double dt, bool enable_cuda,
                                                                         unsigned int cuda_thread_per_block)
    : dt_(dt), enable_cuda_(enable_cuda), cuda_thread_per_block_(cuda_thread_per_block) {}

ElastodynamicsCQBEMMatrixCollection::~ElastodynamicsCQBEMMatrixCollection() {
    free_resident();
    free_assembled();
}

void ElastodynamicsCQBEMMatrixCollection::free_assembled() const {
    for (double *p : {assembled_G_, assembled_H_, assembled_B_})
        if (p) (void)hipFree(p);
    assembled_G_ = assembled_H_ = assembled_B_ = nullptr;
}

void ElastodynamicsCQBEMMatrixCollection::free_resident() {
    for (double *p : {resident_.G, resident_.H, resident_.B, resident_.u_p, resident_.H_inv})
        if (p) (void)hipFree(p);
    for (auto &m : resident_.Gc) m.release();
    for (auto &m : resident_.Hc) m.release();
    resident_.u_p_c.release(), resident_.H_inv_c.release();
    resident_ = Resident{};
    resident_ready_ = false;
}

const ElastodynamicsCQBEMMatrixCollection::Resident &ElastodynamicsCQBEMMatrixCollection::resident() const {
    std::lock_guard<std::mutex> lk(resident_mu_);
    if (resident_ready_) return resident_;
    const size_t L = G_matrices.size();
    const int64_t n3 = 3 * num_vertices;
    if (L == 0 || (int64_t)u_p.data.size() != n3 * n3) throw std::runtime_error("collection not initialised");
    std::string err;
    Resident r;
    r.ctx = context(err);
    if (!r.ctx) throw std::runtime_error(err);
    auto fail_up = [&](const char *what) {
        for (double *p : {r.G, r.H, r.B, r.u_p, r.H_inv})
            if (p) (void)hipFree(p);
        for (auto &m : r.Gc) m.release();
        for (auto &m : r.Hc) m.release();
        r.u_p_c.release(), r.H_inv_c.release();
        throw std::runtime_error(std::string("resident matrices: ") + what);
    };
    const size_t mat = sizeof(double) * (size_t)(n3 * n3);
    if (use_compressed_matrices) {
        // one dense scratch matrix at a time, compressed on the device (CompressedMatrix(mat, ratio) keeps
        // (int)(mat.size() * ratio) coefficients, wavelets.hpp:443-445).  With a cache folder, each compressed
        // matrix is read from its .cpmat file when present and written there otherwise, under the reference's
        // names (matrix_collection.cpp:97-134,229-271).
        const int64_t keep = (int64_t)(int)((double)(n3 * n3) * compression_ratio);
        const std::string tag = "_" + std::to_string(compression_ratio) + ".cpmat";
        double *scratch = nullptr;
        if (hipMalloc(&scratch, mat) != hipSuccess) fail_up("hipMalloc failed");
        // fill() puts the dense matrix somewhere on the device and returns it with the factor for its kept values
        using Source = std::pair<const double *, double>;
        auto get = [&](DeviceCsr &out, const std::string &name, const std::function<Source()> &fill) {
            const std::string path = cache_folder.empty() ? "" : cache_folder + "/" + name + tag;
            if (!path.empty() && load_cpmat(out, n3, path)) {
                compressed_loaded++;
                return;
            }
            const Source src = fill();
            wavelet_compress(src.first, n3, keep, src.second, out, nullptr);
            if (!path.empty() && !save_cpmat(out, path)) throw std::runtime_error("cannot write " + path);
        };
        auto upload_host = [&](const Dense &m) {
            if (hipMemcpy(scratch, m.data.data(), mat, hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("upload failed");
            return Source{scratch, 1.0};
        };
        try {
            r.Gc.resize(L), r.Hc.resize(L);
            for (size_t n = 0; n < L; n++) {
                // init's device copies hold H_n un-negated: compressed with factor -1, the kept set and values
                // are those of the negated host matrix (negation commutes exactly with the Haar steps)
                get(r.Hc[n], "H_" + std::to_string(n), [&] {
                    return assembled_H_ ? Source{assembled_H_ + n * n3 * n3, -1.0} : upload_host(H_matrices[n]);
                });
                get(r.Gc[n], "G_" + std::to_string(n), [&] {
                    return assembled_G_ ? Source{assembled_G_ + n * n3 * n3, 1.0} : upload_host(G_matrices[n]);
                });
            }
            get(r.u_p_c, update_local_frame ? "u_p" : "u_p_no_frame_update", [&] { return upload_host(u_p); });
            get(r.H_inv_c, "H_inv", [&] {
                if (!H_inv.inverse_to_device(scratch, err)) throw std::runtime_error(err);
                return Source{scratch, 1.0};
            });
            r.vlen = 2 * wavelet_padded(n3) - 2;
        } catch (const std::exception &e) {
            (void)hipFree(scratch);
            fail_up(e.what());
        }
        (void)hipFree(scratch);
        if (assembled_B_) {
            r.B = assembled_B_;
            assembled_B_ = nullptr;
        } else {
            if (hipMalloc(&r.B, sizeof(double) * (size_t)(n3 * 6) * L) != hipSuccess) fail_up("hipMalloc failed");
            for (size_t n = 0; n < L; n++)
                if (hipMemcpy(r.B + n * n3 * 6, B_matrices[n].data.data(), sizeof(double) * n3 * 6, hipMemcpyHostToDevice) != hipSuccess)
                    fail_up("upload failed");
        }
        free_assembled();  // the dense G_n and H_n are not stepped on
        resident_ = std::move(r);
        resident_ready_ = true;
        return resident_;
    }
    if (assembled_G_ && assembled_H_ && assembled_B_) {  // init's device copies, in the layout cqbem_history reads
        r.G = assembled_G_, r.H = assembled_H_, r.B = assembled_B_;
        assembled_G_ = assembled_H_ = assembled_B_ = nullptr;
        if (hipMalloc(&r.u_p, mat) != hipSuccess || hipMalloc(&r.H_inv, mat) != hipSuccess) fail_up("hipMalloc failed");
        if (hipMemcpy(r.u_p, u_p.data.data(), mat, hipMemcpyHostToDevice) != hipSuccess) fail_up("upload failed");
        if (!H_inv.inverse_to_device(r.H_inv, err)) fail_up(err.c_str());
        resident_ = std::move(r);
        resident_ready_ = true;
        return resident_;
    }
    if (hipMalloc(&r.G, mat * L) != hipSuccess || hipMalloc(&r.H, mat * L) != hipSuccess ||
        hipMalloc(&r.B, sizeof(double) * (size_t)(n3 * 6) * L) != hipSuccess || hipMalloc(&r.u_p, mat) != hipSuccess ||
        hipMalloc(&r.H_inv, mat) != hipSuccess)
        fail_up("hipMalloc failed");
    std::vector<double> h((size_t)(n3 * n3));
    for (size_t n = 0; n < L; n++) {
        parallel_for((int64_t)h.size(), [&](int64_t a, int64_t b) {  // back to the assembled sign
            for (int64_t e = a; e < b; e++) h[e] = -H_matrices[n].data[e];
        });
        if (hipMemcpy(r.G + n * n3 * n3, G_matrices[n].data.data(), mat, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(r.H + n * n3 * n3, h.data(), mat, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(r.B + n * n3 * 6, B_matrices[n].data.data(), sizeof(double) * n3 * 6, hipMemcpyHostToDevice) != hipSuccess)
            fail_up("upload failed");
    }
    if (hipMemcpy(r.u_p, u_p.data.data(), mat, hipMemcpyHostToDevice) != hipSuccess) fail_up("upload failed");
    if (!H_inv.inverse_to_device(r.H_inv, err)) fail_up(err.c_str());
    resident_ = std::move(r);
    resident_ready_ = true;
    return resident_;
}

bool ElastodynamicsCQBEMMatrixCollection::fail(const std::string &msg) {
    error_ = msg;
    return false;
}

bool prepare_mesh(const Json &m, std::vector<double> &v, std::vector<int32_t> &f, std::vector<int32_t> &vmap,
                  std::string &err) {
    vmap.clear();
    if (!load_obj(m["file"].string(), v, f, err)) return false;
    // unused vertices dropped, the survivors keeping their relative order (lines 75-103)
    {
        const int64_t nv0 = (int64_t)(v.size() / 3);
        std::vector<int32_t> used(nv0, -1);
        for (int32_t x : f) used[x] = 0;
        int32_t nv1 = 0;
        for (int64_t i = 0; i < nv0; i++)
            if (used[i] == 0) used[i] = nv1++;
        if (nv1 != nv0) {
            std::vector<double> v1(3 * (size_t)nv1);
            for (int64_t i = 0; i < nv0; i++)
                if (used[i] >= 0)
                    for (int k = 0; k < 3; k++) v1[3 * (size_t)used[i] + k] = v[3 * i + k];
            for (int32_t &x : f) x = used[x];
            v.swap(v1);
        }
    }
    if (m.contains("scale"))
        for (double &x : v) x *= m["scale"].number();
    // subdivision (lines 109-140): subdivision_level rounds; an unknown scheme name means Upsample
    if (m.contains("subdivision_level") && m["subdivision_level"].number() > 0) {
        const std::string name = m.contains("subdivision_scheme") ? m["subdivision_scheme"].string() : "Upsample";
        const int scheme = name == "Loop" ? CQBEM_SUBDIV_LOOP
                           : name == "FalseBarycentric" ? CQBEM_SUBDIV_FALSE_BARYCENTRIC
                                                        : CQBEM_SUBDIV_UPSAMPLE;
        if (subdivide_mesh(v, f, scheme, (int)m["subdivision_level"].number(), err) != CQBEM_OK)
            return err = "mesh subdivision: " + err, false;
    }
    // decimation (lines 143-151) to (Eigen::Index)(faces * decimation_scale) faces
    if (m.contains("decimation_scale") && m["decimation_scale"].number() < 1.0) {
        const int64_t target = (int64_t)((double)(f.size() / 3) * m["decimation_scale"].number());
        if (decimate_mesh(v, f, target, err) != CQBEM_OK) return err = "mesh decimation: " + err, false;
    }
    // lines 158-196: a build with OpenCCL reorders the vertices (default on); without OpenCCL the reference
    // ignores the flag.  The cache key does not record the order, so an explicit request is refused rather
    // than silently producing (or reading cached) matrices in another vertex order.
    if (m.contains("enable_vertex_reordering") && m["enable_vertex_reordering"].boolean())
        return err = "mesh.enable_vertex_reordering = true is not supported: reorder the OBJ beforehand (the shipped "
                     "scenes ship pre-reordered meshes and set it to false)",
               false;
    if (m.contains("flip_normal") && m["flip_normal"].boolean())  // line 152
        for (size_t t = 0; t < f.size(); t += 3) std::swap(f[t + 1], f[t + 2]);
    if (m.contains("enable_traction_discontinuity") && m["enable_traction_discontinuity"].boolean()) {
        // lines 68-71, 203-365
        const double deg = m.contains("sharp_edge_threshold_degree") ? m["sharp_edge_threshold_degree"].number() : 20.0;
        std::vector<double> vs;
        std::vector<int32_t> fs;
        if (split_traction_discontinuity(v.data(), (int)(v.size() / 3), f.data(), (int)(f.size() / 3), deg, vs, fs, vmap,
                                         err) != CQBEM_OK)
            return err = "traction-discontinuity split: " + err, false;
        v.swap(vs), f.swap(fs);
    }
    return true;
}

bool ElastodynamicsCQBEMMatrixCollection::init(const Json &config) {
    try {
        std::vector<double> v;
        std::vector<int32_t> f, vm;
        std::string err;
        if (!prepare_mesh(config["mesh"], v, f, vm, err)) return fail(err);
        return init_mesh(config, std::move(v), std::move(f), std::move(vm));
    } catch (const std::exception &e) {
        return fail(e.what());
    }
}

bool ElastodynamicsCQBEMMatrixCollection::init(const Json &config, const std::vector<double> &v,
                                               const std::vector<int32_t> &f) {
    return init_mesh(config, v, f, {});
}

bool ElastodynamicsCQBEMMatrixCollection::init_mesh(const Json &config, std::vector<double> v, std::vector<int32_t> f,
                                                    std::vector<int32_t> vmap) {
    try {
        V = std::move(v), F = std::move(f), vertex_map_inverse = std::move(vmap);
        num_vertices = (int64_t)(V.size() / 3);
        return run(config);
    } catch (const std::exception &e) {
        return fail(e.what());
    }
}

bool ElastodynamicsCQBEMMatrixCollection::run(const Json &config) {
    if (!enable_cuda_) return fail("the CQBEM precompute of this build runs on the GPU only (enable_cuda = false)");
    const Json &mat = config["material"], &sim = config["simulation"];
    rho = mat["rho"].number();
    nu = mat["nu"].number();
    mu = mat["mu"].number();
    // centre on the volume centroid (physics_base_matrix_collection.cpp:41-45)
    {
        double vol = 0, c[3] = {0, 0, 0};
        for (size_t t = 0; t < F.size(); t += 3) {
            const double *a = &V[3 * F[t]], *b = &V[3 * F[t + 1]], *d = &V[3 * F[t + 2]];
            const double tv = (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                               a[2] * (b[0] * d[1] - b[1] * d[0])) / 6.0;
            vol += tv;
            for (int k = 0; k < 3; k++) c[k] += tv * (a[k] + b[k] + d[k]) / 4.0;
        }
        if (!(vol > 0)) return fail("mesh must be closed and oriented outward (non-positive volume)");
        for (int k = 0; k < 3; k++) original_cm[k] = c[k] / vol, cm[k] = 0.0;
        for (int64_t i = 0; i < num_vertices; i++)
            for (int k = 0; k < 3; k++) V[3 * i + k] -= original_cm[k];
        mass = vol * rho;
    }
    if (dt_ != sim["dt"].number()) return fail("dt for simulator and elastodynamics BEM need to be the same.");
    const std::string ms = sim["multistep_method"].string();
    if (ms == "BDF1")
        multistep_method = BDF1;
    else if (ms == "BDF2")
        multistep_method = BDF2;
    else
        return fail("Unsupported multistep method detected. Valid options are \"BDF1\"/\"BDF2\"");
    use_galerkin = sim.contains("use_galerkin") && sim["use_galerkin"].boolean();
    update_local_frame = !sim.contains("update_local_frame") || sim["update_local_frame"].boolean();
    compression_ratio = sim.contains("compression_ratio") ? sim["compression_ratio"].number() : 0.0;
    use_compressed_matrices = compression_ratio != 0.0;
    const int order = (int)sim["gaussian_quadrature_order"].number();
    const int sub = (int)sim["quadrature_subdivision"].number();
    // material constants and history length (matrix_collection.cpp:59-69)
    const double lambda = 2 * nu * mu / (1. - 2. * nu);
    c1 = std::sqrt((lambda + 2. * mu) / rho);
    c2 = std::sqrt(mu / rho);
    if (sim.contains("cq_length")) {
        max_time_history = (size_t)sim["cq_length"].number() - 1;
    } else {
        double maxd = 0.0;
        for (int64_t j = 0; j < num_vertices; j++)
            for (int64_t i = 0; i < num_vertices; i++) {
                const double dx = V[3 * i] - V[3 * j], dy = V[3 * i + 1] - V[3 * j + 1], dz = V[3 * i + 2] - V[3 * j + 2];
                maxd = std::max(maxd, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        max_time_history = (size_t)std::floor(maxd / (c2 * dt_) + 2);
    }
    R = sim.contains("cq_radius") ? sim["cq_radius"].number() : std::pow(1e-10, 1. / (2. * max_time_history));
    const size_t L = max_time_history + 1;

    // the reference's matrix cache: <coeffs_folder_path>/elastodynamicsCQBEM_<key>/{G,H}_<n>.mat, H stored
    // un-negated, config.json beside them (matrix_collection.cpp:41-57,77-94,135-143)
    std::string folder;
    if (sim.contains("coeffs_folder_path")) {
        Json stripped;
        folder = sim["coeffs_folder_path"].string() + "/elastodynamicsCQBEM_" + cache_key(config, &stripped);
        std::filesystem::create_directories(folder);
        std::ofstream(folder + "/config.json") << stripped.dump(4);
    }
    cache_folder = folder;
    const int64_t n3 = 3 * num_vertices;
    G_matrices.assign(L, Dense(n3, n3));
    H_matrices.assign(L, Dense(n3, n3));
    B_matrices.assign(L, Dense(n3, 6));
    loaded_from_cache = !folder.empty();
    for (size_t n = 0; n < L && loaded_from_cache; n++)
        loaded_from_cache = std::filesystem::exists(folder + "/H_" + std::to_string(n) + ".mat") &&
                            std::filesystem::exists(folder + "/G_" + std::to_string(n) + ".mat");
    if (loaded_from_cache)
        for (size_t n = 0; n < L; n++) {
            if (!load_matrix(H_matrices[n], folder + "/H_" + std::to_string(n) + ".mat") ||
                !load_matrix(G_matrices[n], folder + "/G_" + std::to_string(n) + ".mat") || G_matrices[n].rows != n3 ||
                G_matrices[n].cols != n3 || H_matrices[n].rows != n3 || H_matrices[n].cols != n3)
                return fail("cached matrices in " + folder + " do not match this mesh");
            for (double &x : H_matrices[n].data) x = -x;  // matrix_collection.cpp:140
        }

    std::string err;
    cqbem_context *ctx = context(err);
    if (!ctx) return fail(err);
    free_resident();
    free_assembled();
    // rocSOLVER's first-use cost (its workspace allocations included), paid beside the download of the
    // assembled matrices instead of inside post_process().  It starts only after cqbem_assemble has returned,
    // so its allocations never compete with the assembly's, which size the spectra slabs from free memory.
    struct Joined {
        std::thread t;
        void start() {
            if (!t.joinable()) t = std::thread(warm_up_lu);
        }
        void join() {
            if (t.joinable()) t.join();
        }
        ~Joined() { join(); }
    } warm;
    cqbem_mesh mesh{V.data(), num_vertices, F.data(), (int64_t)(F.size() / 3),
                    vertex_map_inverse.empty() ? nullptr : vertex_map_inverse.data()};
    // with cached G and H only B is assembled (the reference computes B on every init, :146-177)
    cqbem_params prm{c1, c2, rho, nu, dt_, R, (int32_t)L, multistep_method == BDF1 ? CQBEM_BDF1 : CQBEM_BDF2,
                     order, sub, use_galerkin ? 1 : 0, CQBEM_FP64,
                     loaded_from_cache ? (CQBEM_B_TRANS | CQBEM_B_ANGULAR) : CQBEM_ALL};
    const auto t0 = std::chrono::steady_clock::now();
    // Assembled into device buffers of the collection when they can be allocated: the host matrices are copied
    // from there and the time stepping keeps them (resident()).  The library's own host-output path needs the
    // same device memory, so a failed allocation means that path too would hold the outputs only while staging.
    const size_t mbytes = sizeof(double) * (size_t)(n3 * n3), bmat = sizeof(double) * (size_t)(n3 * 6);
    {
        double *dg = nullptr, *dh = nullptr, *db = nullptr;
        const bool ok = (loaded_from_cache || (hipMalloc(&dg, mbytes * L) == hipSuccess && hipMalloc(&dh, mbytes * L) == hipSuccess)) &&
                        hipMalloc(&db, bmat * L) == hipSuccess;
        if (ok) {
            assembled_G_ = dg, assembled_H_ = dh, assembled_B_ = db;
        } else {
            for (double *q : {dg, dh, db})
                if (q) (void)hipFree(q);
            (void)hipGetLastError();
        }
    }
    if (assembled_B_) {
        cqbem_outputs out{assembled_G_, assembled_H_, nullptr, nullptr, assembled_B_, 1};
        if (cqbem_assemble(ctx, &mesh, &prm, &out) != CQBEM_OK) {
            free_assembled();
            return fail(cqbem_last_error(ctx));
        }
        warm.start();
        DownloadStager stager;
        for (size_t n = 0; n < L; n++) {
            if (!loaded_from_cache) {
                double *gd = G_matrices[n].data.data(), *hd = H_matrices[n].data.data();
                const double *gs = assembled_G_ + n * n3 * n3, *hs = assembled_H_ + n * n3 * n3;
                // H un-negated first when the cache folder is written (the .mat files hold the assembled sign)
                const double hsign = folder.empty() ? -1.0 : 1.0;  // matrix_collection.cpp:140
                const bool copied = stager.ok() ? stager.copy(gd, gs, n3 * n3, 1.0) && stager.copy(hd, hs, n3 * n3, hsign)
                                                : hipMemcpy(gd, gs, mbytes, hipMemcpyDeviceToHost) == hipSuccess &&
                                                      hipMemcpy(hd, hs, mbytes, hipMemcpyDeviceToHost) == hipSuccess;
                if (!copied) {
                    free_assembled();
                    return fail("device-to-host copy of the matrices failed");
                }
                if (!folder.empty() || !stager.ok()) {
                    if (!folder.empty() && (!save_matrix(H_matrices[n], folder + "/H_" + std::to_string(n) + ".mat") ||
                                            !save_matrix(G_matrices[n], folder + "/G_" + std::to_string(n) + ".mat")))
                        return fail("cannot write the matrix cache in " + folder);
                    parallel_for(n3 * n3, [&](int64_t a, int64_t e1) {
                        for (int64_t e = a; e < e1; e++) hd[e] = -hd[e];  // matrix_collection.cpp:140
                    });
                }
            }
            if (hipMemcpy(B_matrices[n].data.data(), assembled_B_ + n * n3 * 6, bmat, hipMemcpyDeviceToHost) != hipSuccess) {
                free_assembled();
                return fail("device-to-host copy of B failed");
            }
        }
        if (loaded_from_cache) free_assembled();  // G and H come from the files: resident() uploads them
        last_assembly_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        (void)cuda_thread_per_block_;  // launch shapes are chosen by the library
        warm.join();
        const auto t1 = std::chrono::steady_clock::now();
        if (!post_process()) return false;
        last_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        return true;
    }
    std::vector<double> g, h, b((size_t)L * n3 * 6);
    if (!loaded_from_cache) g.resize((size_t)L * n3 * n3), h.resize((size_t)L * n3 * n3);
    cqbem_outputs out{loaded_from_cache ? nullptr : g.data(), loaded_from_cache ? nullptr : h.data(), nullptr, nullptr,
                      b.data(), 0};
    if (cqbem_assemble(ctx, &mesh, &prm, &out) != CQBEM_OK) return fail(cqbem_last_error(ctx));
    warm.start();
    for (size_t n = 0; n < L; n++) {
        if (!loaded_from_cache) {
            const double *gs = g.data() + n * n3 * n3, *hs = h.data() + n * n3 * n3;
            double *gd = G_matrices[n].data.data(), *hd = H_matrices[n].data.data();
            if (!folder.empty()) {
                std::copy(hs, hs + n3 * n3, hd);
                std::copy(gs, gs + n3 * n3, gd);
                if (!save_matrix(H_matrices[n], folder + "/H_" + std::to_string(n) + ".mat") ||
                    !save_matrix(G_matrices[n], folder + "/G_" + std::to_string(n) + ".mat"))
                    return fail("cannot write the matrix cache in " + folder);
            }
            parallel_for(n3 * n3, [&](int64_t a, int64_t e1) {
                std::copy(gs + a, gs + e1, gd + a);
                for (int64_t e = a; e < e1; e++) hd[e] = -hs[e];  // matrix_collection.cpp:140
            });
        }
        std::copy(b.begin() + n * n3 * 6, b.begin() + (n + 1) * n3 * 6, B_matrices[n].data.begin());
    }
    last_assembly_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    (void)cuda_thread_per_block_;  // launch shapes are chosen by the library
    warm.join();
    const auto t1 = std::chrono::steady_clock::now();
    if (!post_process()) return false;
    last_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    return true;
}

// What init() derives from G_0, H_0 and B_0 for the time stepping: the external acceleration matrices
// (physics_base_matrix_collection.cpp:40-50,374-436), G' and the least-squares / LU solves
// (matrix_collection.cpp:174-215), the rotation-removal weights (:307-322) and u_p's diagonal blocks
// (:217-227).  The LU of H_0 and its solves run on the GPU.
bool ElastodynamicsCQBEMMatrixCollection::post_process() {
    const int64_t nv = num_vertices, n3 = 3 * nv, nf = (int64_t)(F.size() / 3);
    auto tp = std::chrono::steady_clock::now();
    auto phase = [&](int k) {
        const auto t = std::chrono::steady_clock::now();
        post_process_phase_ms[k] = std::chrono::duration<double, std::milli>(t - tp).count();
        tp = t;
    };
    auto vtx = [&](int64_t v, int k) { return V[3 * v + k]; };
    std::vector<double> A2(nf), N(3 * nf);  // doubled areas, unit normals
    for (int64_t f = 0; f < nf; f++) {
        const int32_t *t = &F[3 * f];
        double e1[3], e2[3], c[3];
        for (int k = 0; k < 3; k++) e1[k] = vtx(t[1], k) - vtx(t[0], k), e2[k] = vtx(t[2], k) - vtx(t[0], k);
        c[0] = e1[1] * e2[2] - e1[2] * e2[1], c[1] = e1[2] * e2[0] - e1[0] * e2[2], c[2] = e1[0] * e2[1] - e1[1] * e2[0];
        A2[f] = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int k = 0; k < 3; k++) N[3 * f + k] = c[k] / A2[f];
    }
    // inertia about the volume centroid of the centred mesh, order-4 triangle rule
    double cmc[3] = {0, 0, 0}, vol = 0;
    for (int64_t f = 0; f < nf; f++) {
        const double *a = &V[3 * F[3 * f]], *b = &V[3 * F[3 * f + 1]], *d = &V[3 * F[3 * f + 2]];
        const double tv = (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                           a[2] * (b[0] * d[1] - b[1] * d[0])) / 6.0;
        vol += tv;
        for (int k = 0; k < 3; k++) cmc[k] += tv * (a[k] + b[k] + d[k]) / 4.0;
    }
    for (int k = 0; k < 3; k++) cmc[k] /= vol;
    std::fill(I, I + 9, 0.0);
    const auto rule = tri_rule(4);
    for (int64_t f = 0; f < nf; f++) {
        const double *n = &N[3 * f];
        for (const TriPoint &q : rule) {
            double r[3];
            for (int k = 0; k < 3; k++)
                r[k] = q.b[0] * vtx(F[3 * f], k) + q.b[1] * vtx(F[3 * f + 1], k) + q.b[2] * vtx(F[3 * f + 2], k) - cmc[k];
            const double w = q.w * 0.5 * A2[f];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
                    const double v = i == j ? (rho / 3.) * (r[i1] * r[i1] * r[i1] * n[i1] + r[i2] * r[i2] * r[i2] * n[i2])
                                            : (-rho / 4.) * (r[j] * r[i] * r[i] * n[i] + r[i] * r[j] * r[j] * n[j]);
                    I[3 * i + j] += w * v;
                }
        }
    }
    // external acceleration matrices (3 x 3 nv)
    translational_acc_mat = Dense(3, n3);
    rotational_acc_mat = Dense(3, n3);
    Dense rot_raw(3, n3);
    for (int64_t f = 0; f < nf; f++) {
        double sum[3];
        for (int k = 0; k < 3; k++) sum[k] = vtx(F[3 * f], k) + vtx(F[3 * f + 1], k) + vtx(F[3 * f + 2], k);
        for (int c = 0; c < 3; c++) {
            const int64_t j = F[3 * f + c];
            for (int k = 0; k < 3; k++) translational_acc_mat(k, 3 * j + k) += A2[f] / 6.;
            double v[3];
            for (int k = 0; k < 3; k++) v[k] = A2[f] / 6. * 0.25 * (sum[k] + vtx(j, k));
            // cross-product matrix [v]_x
            const double X[3][3] = {{0, -v[2], v[1]}, {v[2], 0, -v[0]}, {-v[1], v[0], 0}};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) rot_raw(a, 3 * j + b) += X[a][b];
        }
    }
    for (double &x : translational_acc_mat.data) x *= 1. / mass;
    {
        // I^-1 by the adjugate
        const double *m = I;
        const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
        if (!(std::fabs(det) > 0)) return fail("singular inertia tensor");
        const double inv[9] = {(m[4] * m[8] - m[5] * m[7]) / det, (m[2] * m[7] - m[1] * m[8]) / det, (m[1] * m[5] - m[2] * m[4]) / det,
                               (m[5] * m[6] - m[3] * m[8]) / det, (m[0] * m[8] - m[2] * m[6]) / det, (m[2] * m[3] - m[0] * m[5]) / det,
                               (m[3] * m[7] - m[4] * m[6]) / det, (m[1] * m[6] - m[0] * m[7]) / det, (m[0] * m[4] - m[1] * m[3]) / det};
        for (int a = 0; a < 3; a++)
            for (int64_t c = 0; c < n3; c++)
                rotational_acc_mat(a, c) = inv[3 * a] * rot_raw(0, c) + inv[3 * a + 1] * rot_raw(1, c) + inv[3 * a + 2] * rot_raw(2, c);
    }
    phase(0);
    // B_0 halves, H_0 (un-negated) and G'
    std::vector<double> Bt((size_t)(n3 * 3)), Be((size_t)(n3 * 3));
    for (int64_t r = 0; r < n3; r++)
        for (int k = 0; k < 3; k++) Bt[3 * r + k] = B_matrices[0](r, k), Be[3 * r + k] = B_matrices[0](r, 3 + k);
    std::vector<double> H0(H_matrices[0].data.size());
    parallel_for((int64_t)H0.size(), [&](int64_t a, int64_t b) {
        for (int64_t e = a; e < b; e++) H0[e] = -H_matrices[0].data[e];
    });
    // u_p becomes G' = G_0 (+ B_trans T + B_euler R), then H_0^-1 G'; one threaded pass over the rows
    u_p = Dense(n3, n3);
    parallel_for(n3, [&](int64_t r0, int64_t r1) {
        for (int64_t r = r0; r < r1; r++) {
            const double *g0 = &G_matrices[0].data[(size_t)(r * n3)];
            double *up = &u_p.data[(size_t)(r * n3)];
            if (!update_local_frame) {
                std::memcpy(up, g0, sizeof(double) * (size_t)n3);
                continue;
            }
            for (int64_t c = 0; c < n3; c++) {
                double v = 0;
                for (int k = 0; k < 3; k++)
                    v += Bt[3 * r + k] * translational_acc_mat(k, c) + Be[3 * r + k] * rotational_acc_mat(k, c);
                up[c] = g0[c] + v;
            }
        }
    });
    phase(1);
    std::string err;
    if (!B_trans_inv.compute(Bt, n3, 3, err) || !B_euler_inv.compute(Be, n3, 3, err)) return fail(err);
    auto as3 = [&](std::vector<double> x, int64_t cols) {
        Dense d(3, cols);
        d.data = std::move(x);
        return d;
    };
    B_trans_inv_H = as3(B_trans_inv.solve(H0.data(), n3), n3);
    B_euler_inv_H = as3(B_euler_inv.solve(H0.data(), n3), n3);
    B_trans_inv_G_prime = as3(B_trans_inv.solve(u_p.data.data(), n3), n3);
    B_euler_inv_G_prime = as3(B_euler_inv.solve(u_p.data.data(), n3), n3);
    B_euler_inv_B_trans = as3(B_euler_inv.solve(Bt.data(), 3), 3);
    // rotation-removal weights
    rotation_removal_weights.assign(nv, 0.0);
    for (int64_t f = 0; f < nf; f++) {
        double yn = 0;
        for (int k = 0; k < 3; k++) yn += vtx(F[3 * f], k) * N[3 * f + k];
        const double val = (1. / 6.) * A2[f] * yn;
        for (int c = 0; c < 3; c++) rotation_removal_weights[F[3 * f + c]] += val;
    }
    phase(2);
    // LU of H_0 on the GPU and its solves
    if (!H_inv.factor(n3, H0.data(), err)) return fail(err);
    H0.clear(), H0.shrink_to_fit();
    phase(3);
    H_inv_B_trans = Dense(n3, 3), H_inv_B_euler = Dense(n3, 3);
    H_inv_B_trans.data = Bt, H_inv_B_euler.data = Be;
    if (!H_inv.solve(u_p.data.data(), n3, err) || !H_inv.solve(H_inv_B_trans.data.data(), 3, err) ||
        !H_inv.solve(H_inv_B_euler.data.data(), 3, err))
        return fail(err);
    u_p_diag = Dense(nv, 9);
    for (int64_t v = 0; v < nv; v++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) u_p_diag(v, 3 * a + b) = u_p(3 * v + a, 3 * v + b);
    phase(4);
    return true;
}

}  // namespace cqbem
