// service.cc -- reapers_service: one computation per invocation, the request
// a JSON object on standard input and the reply a JSON object on standard
// output, with the reference's protocol (tests/service.cc:21-240):
//
//   {"request": R, "params": {"N": n, "sparsity": k, "regularize": b, ...}}
//
//   R = "gen-ham"           {N, ham: [{bits, coeff}], hamstr, mat}
//       "get-eigensys"      the above and {eigenvals, eigenvecs}
//       "evolve-state"      (params t, beta, exdiag) the above (gen-ham's)
//                           and {init-state, final-state}: {len, vec}
//       "matexp"            (params t, beta) gen-ham's and {res} =
//                           exp(-(beta + i t) H)
//       "get-ham-sparsity"  {sparsityL, sparsityR} of the parity-block form
//
// The Hamiltonian is SYK(N, sparsity).gen_ham on N/2 sites in FP64 (the
// non-block form; get-ham-sparsity uses the block form); "coeff" is the
// real part of the coefficient in the standard sigma_y convention, matrices
// are row-major lists of [re, im] pairs, vectors lists of [re, im].  A
// "seed" parameter (not in the reference, which seeds from
// std::random_device) makes a reply reproducible; the random initial state
// of evolve-state then uses seed + 1.  Errors go to standard error with
// exit status 1.
#include "reapers/reapers.h"

#include <cctype>
#include <cstdio>
#include <iostream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <vector>

using namespace REAPERS;
using namespace REAPERS::Model;
using namespace REAPERS::EvolutionAlgorithm;

namespace {

// ---- a small JSON value and reader ----------------------------------------

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json &at(const std::string &k) const {
        auto it = obj.find(k);
        if (kind != Object || it == obj.end()) throw std::runtime_error("missing key \"" + k + "\"");
        return it->second;
    }
    bool has(const std::string &k) const { return kind == Object && obj.count(k); }
    double number() const {
        if (kind != Number) throw std::runtime_error("expected a number");
        return num;
    }
    bool boolean() const {
        if (kind != Bool) throw std::runtime_error("expected true or false");
        return b;
    }
    const std::string &string() const {
        if (kind != String) throw std::runtime_error("expected a string");
        return str;
    }
};

class Reader {
    const std::string &s;
    size_t i = 0;

    [[noreturn]] void fail(const char *what) {
        throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool lit(const char *w) {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    std::string parse_string() {
        if (s[i] != '"') fail("expected a string");
        i++;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("bad escape");
            char e = s[i++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                if (i + 4 > s.size()) fail("bad \\u escape");
                unsigned v = std::stoul(s.substr(i, 4), nullptr, 16);
                i += 4;
                if (v < 0x80) out += (char)v;
                else if (v < 0x800) {
                    out += (char)(0xc0 | (v >> 6));
                    out += (char)(0x80 | (v & 0x3f));
                } else {
                    out += (char)(0xe0 | (v >> 12));
                    out += (char)(0x80 | ((v >> 6) & 0x3f));
                    out += (char)(0x80 | (v & 0x3f));
                }
                break;
            }
            default: fail("bad escape");
            }
        }
        if (i >= s.size()) fail("unterminated string");
        i++;
        return out;
    }

public:
    explicit Reader(const std::string &text) : s(text) {}

    Json value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        Json v;
        char c = s[i];
        if (c == '{') {
            v.kind = Json::Object;
            i++;
            ws();
            if (i < s.size() && s[i] == '}') {
                i++;
                return v;
            }
            while (true) {
                ws();
                std::string k = parse_string();
                ws();
                if (i >= s.size() || s[i] != ':') fail("expected ':'");
                i++;
                v.obj[k] = value();
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == '}') { i++; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Json::Array;
            i++;
            ws();
            if (i < s.size() && s[i] == ']') {
                i++;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == ']') { i++; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Json::String;
            v.str = parse_string();
        } else if (lit("true")) {
            v.kind = Json::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Json::Bool;
        } else if (lit("null")) {
        } else {
            const char *p = s.c_str() + i;
            char *end = nullptr;
            v.num = std::strtod(p, &end);
            if (end == p) fail("unexpected character");
            v.kind = Json::Number;
            i += end - p;
        }
        return v;
    }

    void finish() {
        ws();
        if (i != s.size()) fail("trailing characters");
    }
};

// ---- writing --------------------------------------------------------------

std::string num(double x) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", x);
    return buf;
}

std::string quoted(const std::string &s) {
    std::string o = "\"";
    for (char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\n': o += "\\n"; break;
        case '\t': o += "\\t"; break;
        case '\r': o += "\\r"; break;
        default:
            if ((unsigned char)c < 0x20) {
                char b[8];
                std::snprintf(b, sizeof b, "\\u%04x", c);
                o += b;
            } else {
                o += c;
            }
        }
    }
    return o + "\"";
}

struct Obj {
    std::vector<std::pair<std::string, std::string>> kv;
    Obj &add(const std::string &k, const std::string &raw) {
        kv.emplace_back(k, raw);
        return *this;
    }
    std::string str() const {
        std::string o = "{";
        for (size_t i = 0; i < kv.size(); i++) {
            if (i) o += ",";
            o += quoted(kv[i].first) + ":" + kv[i].second;
        }
        return o + "}";
    }
};

template<typename C>
std::string pairs(const C &v) {
    std::string o = "[";
    bool first = true;
    for (const auto &c : v) {
        if (!first) o += ",";
        o += "[" + num(c.real()) + "," + num(c.imag()) + "]";
        first = false;
    }
    return o + "]";
}

// a device matrix as a row-major list of [re, im]
std::string matrix_json(const DevMatrix<double> &m) {
    auto col = m.to_host();   // column-major
    std::vector<complex<double>> rm(col.size());
    for (size_t r = 0; r < m.rows(); r++)
        for (size_t c = 0; c < m.cols(); c++) rm[r * m.cols() + c] = col[r + c * m.rows()];
    return pairs(rm);
}

std::string ham_json(const SumOps<double> &h) {
    std::string o = "[";
    bool first = true;
    for (const auto &op : h) {
        if (!first) o += ",";
        o += Obj().add("bits", std::to_string(op.get_bit_string()))
                 .add("coeff", num(op.coefficient().real()))
                 .str();
        first = false;
    }
    return o + "]";
}

std::string state_json(const State<double> &s) {
    std::vector<complex<double>> v(s.dim());
    s.store(v.data());
    return Obj().add("len", std::to_string(s.spin_chain_length())).add("vec", pairs(v)).str();
}

double sparsity_of(const DevMatrix<double> &m) {
    auto v = m.to_host();
    size_t nz = 0;
    for (auto &c : v)
        if (c != complex<double>{}) nz++;
    return 1.0 - (double)nz / (double)v.size();
}

struct Params {
    int N;
    double sparsity;
    bool regularize;
    uint32_t seed;
    explicit Params(const Json &p) {
        N = (int)p.at("N").number();
        sparsity = p.at("sparsity").number();
        regularize = p.at("regularize").boolean();
        if (p.has("seed")) seed = (uint32_t)p.at("seed").number();
        else seed = std::random_device{}();
    }
};

std::string run(const Json &req)
{
    const std::string &r = req.at("request").string();
    const Json &p = req.at("params");
    Params pr(p);
    std::mt19937 rg(pr.seed);
    if (r == "get-ham-sparsity") {
        SYKBlockForm<double, GPUImpl> syk(pr.N, pr.sparsity);
        auto ham = syk.gen_ham(rg, pr.regularize);
        return Obj().add("sparsityL", num(sparsity_of(ham.LL.get_matrix())))
                    .add("sparsityR", num(sparsity_of(ham.RR.get_matrix())))
                    .str();
    }
    if (r != "gen-ham" && r != "get-eigensys" && r != "evolve-state" && r != "matexp")
        throw std::runtime_error("Unknown request " + r);
    SYK<double, GPUImpl> syk(pr.N, pr.sparsity);
    auto ham = syk.gen_ham(rg, pr.regularize);
    std::stringstream ss;
    ss << ham;
    Obj o;
    o.add("N", std::to_string(pr.N))
     .add("ham", ham_json(ham))
     .add("hamstr", quoted(ss.str()))
     .add("mat", matrix_json(ham.get_matrix()));
    if (r == "get-eigensys") {
        auto es = ham.get_eigensystem(pr.N / 2);
        const auto &vals = std::get<0>(es);
        std::string ev = "[";
        for (size_t i = 0; i < vals.size(); i++) ev += (i ? "," : "") + num(vals[i]);
        o.add("eigenvals", ev + "]").add("eigenvecs", matrix_json(std::get<1>(es)));
    } else if (r == "evolve-state" || r == "matexp") {
        double t = p.at("t").number(), beta = p.at("beta").number();
        if (r == "matexp") {
            o.add("res", matrix_json(ham.matexp(complex<double>{-beta, -t})));
        } else {
            bool exdiag = p.at("exdiag").boolean();
            State<double, GPUImpl> ini(pr.N / 2);
            ini.random_state(p.has("seed") ? (u64)pr.seed + 1 : 0);
            State<double, GPUImpl> fin(ini);
            if (exdiag) fin.evolve<ExactDiagonalization>(ham, t, beta);
            else fin.evolve(ham, t, beta);
            o.add("init-state", state_json(ini)).add("final-state", state_json(fin));
        }
    }
    return o.str();
}

}  // namespace

int main()
{
    try {
        std::string text((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
        Reader rd(text);
        Json req = rd.value();
        rd.finish();
        std::cout << run(req) << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "Program terminated abnormally due to the following error:" << std::endl
                  << e.what() << std::endl;
        return 1;
    }
    return 0;
}
