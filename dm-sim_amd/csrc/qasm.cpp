// qasm.cpp -- OpenQASM 2.0 front-end (see qasm.hpp).
#include "qasm.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <map>
#include <stdexcept>

#include "dmsim_amd.h"
#include "engine.hpp"

namespace dmx {

namespace {

struct QasmError : std::runtime_error {
    QasmError(int line, const std::string& msg)
        : std::runtime_error("line " + std::to_string(line) + ": " + msg)
    {
    }
};

enum TokKind { T_ID, T_NUM, T_STR, T_SYM, T_END };

struct Tok {
    TokKind k;
    std::string s;
    double v = 0;
    int line = 0;
};

std::vector<Tok> lex(const std::string& src)
{
    std::vector<Tok> out;
    int line = 1;
    size_t i = 0;
    const size_t n = src.size();
    while (i < n) {
        const char c = src[i];
        if (c == '\n') {
            line++;
            i++;
        } else if (std::isspace((unsigned char)c)) {
            i++;
        } else if (c == '/' && i + 1 < n && src[i + 1] == '/') {
            while (i < n && src[i] != '\n') i++;
        } else if (c == '/' && i + 1 < n && src[i + 1] == '*') {
            i += 2;
            while (i + 1 < n && !(src[i] == '*' && src[i + 1] == '/')) {
                if (src[i] == '\n') line++;
                i++;
            }
            i += 2;
        } else if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < n && (std::isalnum((unsigned char)src[j]) || src[j] == '_')) j++;
            out.push_back({T_ID, src.substr(i, j - i), 0, line});
            i = j;
        } else if (std::isdigit((unsigned char)c) ||
                   (c == '.' && i + 1 < n && std::isdigit((unsigned char)src[i + 1]))) {
            char* end = nullptr;
            const double v = std::strtod(src.c_str() + i, &end);
            const size_t j = (size_t)(end - src.c_str());
            out.push_back({T_NUM, src.substr(i, j - i), v, line});
            i = j;
        } else if (c == '"') {
            size_t j = i + 1;
            while (j < n && src[j] != '"' && src[j] != '\n') j++;
            if (j >= n || src[j] != '"') throw QasmError(line, "unterminated string");
            out.push_back({T_STR, src.substr(i + 1, j - i - 1), 0, line});
            i = j + 1;
        } else if ((c == '-' && i + 1 < n && src[i + 1] == '>') ||
                   (c == '=' && i + 1 < n && src[i + 1] == '=')) {
            out.push_back({T_SYM, src.substr(i, 2), 0, line});
            i += 2;
        } else if (std::string(";,()[]{}+-*/^").find(c) != std::string::npos) {
            out.push_back({T_SYM, std::string(1, c), 0, line});
            i++;
        } else {
            throw QasmError(line, std::string("unexpected character '") + c + "'");
        }
    }
    out.push_back({T_END, "", 0, line});
    return out;
}

// A built-in gate: the engine op and, per QASM parameter in order, the
// factory slot it fills (0 theta, 1 phi, 2 lambda).
struct Builtin {
    int op;
    std::vector<int> slots;
};

const std::map<std::string, Builtin>& builtins()
{
    static const std::map<std::string, Builtin> t = {
        {"u3", {OP_U3, {0, 1, 2}}},  {"U", {OP_U3, {0, 1, 2}}},  {"u", {OP_U3, {0, 1, 2}}},
        {"u2", {OP_U2, {1, 2}}},     {"u1", {OP_U1, {2}}},        {"p", {OP_U1, {2}}},
        {"cx", {OP_CX, {}}},         {"CX", {OP_CX, {}}},         {"id", {OP_ID, {}}},
        {"x", {OP_X, {}}},           {"y", {OP_Y, {}}},           {"z", {OP_Z, {}}},
        {"h", {OP_H, {}}},           {"s", {OP_S, {}}},           {"sdg", {OP_SDG, {}}},
        {"t", {OP_T, {}}},           {"tdg", {OP_TDG, {}}},       {"rx", {OP_RX, {0}}},
        {"ry", {OP_RY, {0}}},        {"rz", {OP_RZ, {1}}},        {"cz", {OP_CZ, {}}},
        {"cy", {OP_CY, {}}},         {"swap", {OP_SWAP, {}}},     {"ch", {OP_CH, {}}},
        {"ccx", {OP_CCX, {}}},       {"cswap", {OP_CSWAP, {}}},   {"crx", {OP_CRX, {2}}},
        {"cry", {OP_CRY, {2}}},      {"crz", {OP_CRZ, {2}}},      {"cu1", {OP_CU1, {2}}},
        {"cp", {OP_CU1, {2}}},       {"cu3", {OP_CU3, {0, 1, 2}}}, {"rxx", {OP_RXX, {0}}},
        {"rzz", {OP_RZZ, {0}}},      {"ryy", {OP_RYY, {0}}},      {"rccx", {OP_RCCX, {}}},
        {"rc3x", {OP_RC3X, {}}},     {"c3x", {OP_C3X, {}}},       {"c3sqrtx", {OP_C3SQRTX, {}}},
        {"c4x", {OP_C4X, {}}},       {"r", {OP_R, {0}}},          {"srn", {OP_SRN, {}}},
        {"w", {OP_W, {}}},
    };
    return t;
}

// Expanded gate count limit (nested user gates can multiply a short text).
constexpr size_t kMaxGates = 1u << 22;
// Gate applications, user-defined ones included: nested definitions whose
// bodies emit nothing (empty or barrier-only) still multiply the work.
constexpr size_t kMaxApplications = 1u << 24;
// Nesting of expressions, unary signs and powers in one expression (the
// recursive descent must not exhaust the host stack).
constexpr int kMaxExprDepth = 1024;

struct GateDef {
    std::vector<std::string> params;
    std::vector<std::string> qargs;
    size_t body_begin = 0, body_end = 0;  // token range inside the braces
    bool opaque = false;
};

struct Parser {
    std::vector<Tok> toks;
    size_t pos = 0;
    std::map<std::string, std::pair<int, int>> qregs;  // name -> (offset, size)
    std::map<std::string, int> cregs;
    std::map<std::string, GateDef> defs;
    QasmProgram* prog = nullptr;
    size_t applications = 0;
    int expr_depth = 0;

    // Counts one level of expression nesting for the lifetime of the object.
    struct Nest {
        Parser& p;
        Nest(Parser& parser, int line) : p(parser)
        {
            if (++p.expr_depth > kMaxExprDepth) {
                p.expr_depth = 0;
                throw QasmError(line, "expression nests more than " +
                                          std::to_string(kMaxExprDepth) + " levels deep");
            }
        }
        ~Nest()
        {
            if (p.expr_depth > 0) p.expr_depth--;
        }
    };

    const Tok& peek() const { return toks[pos]; }
    const Tok& next() { return toks[pos++]; }
    bool is_sym(const char* s) const { return peek().k == T_SYM && peek().s == s; }
    bool is_id(const char* s) const { return peek().k == T_ID && peek().s == s; }
    void expect_sym(const char* s)
    {
        if (!is_sym(s))
            throw QasmError(peek().line, std::string("expected '") + s + "', found '" +
                                             peek().s + "'");
        pos++;
    }
    std::string expect_id()
    {
        if (peek().k != T_ID) throw QasmError(peek().line, "expected a name, found '" + peek().s + "'");
        return next().s;
    }
    int expect_int()
    {
        if (peek().k != T_NUM) throw QasmError(peek().line, "expected an integer");
        const Tok& t = next();
        if (t.v != std::floor(t.v) || t.v < 0 || t.v > 1e6)
            throw QasmError(t.line, "expected a non-negative integer, found " + t.s);
        return (int)t.v;
    }
    void skip_statement()
    {
        while (peek().k != T_END && !is_sym(";")) pos++;
        if (is_sym(";")) pos++;
    }

    // ---- expressions (real-valued, with the gate's parameters in env) ----
    using Env = std::map<std::string, double>;
    double expr(const Env& env)
    {
        Nest nest(*this, peek().line);
        double v = term(env);
        while (is_sym("+") || is_sym("-")) {
            const bool add = next().s == "+";
            const double r = term(env);
            v = add ? v + r : v - r;
        }
        return v;
    }
    double term(const Env& env)
    {
        double v = unary(env);
        while (is_sym("*") || is_sym("/")) {
            const bool mul = next().s == "*";
            const double r = unary(env);
            v = mul ? v * r : v / r;
        }
        return v;
    }
    double unary(const Env& env)
    {
        Nest nest(*this, peek().line);
        if (is_sym("-")) {
            pos++;
            return -unary(env);
        }
        if (is_sym("+")) {
            pos++;
            return unary(env);
        }
        const double b = primary(env);
        if (is_sym("^")) {
            pos++;
            return std::pow(b, unary(env));
        }
        return b;
    }
    double primary(const Env& env)
    {
        const Tok& t = next();
        if (t.k == T_NUM) return t.v;
        if (t.k == T_SYM && t.s == "(") {
            const double v = expr(env);
            expect_sym(")");
            return v;
        }
        if (t.k == T_ID) {
            if (t.s == "pi") return 3.14159265358979323846;
            auto it = env.find(t.s);
            if (it != env.end()) return it->second;
            static const std::map<std::string, double (*)(double)> fn = {
                {"sin", [](double x) { return std::sin(x); }},
                {"cos", [](double x) { return std::cos(x); }},
                {"tan", [](double x) { return std::tan(x); }},
                {"exp", [](double x) { return std::exp(x); }},
                {"ln", [](double x) { return std::log(x); }},
                {"sqrt", [](double x) { return std::sqrt(x); }},
            };
            auto f = fn.find(t.s);
            if (f != fn.end()) {
                expect_sym("(");
                const double v = expr(env);
                expect_sym(")");
                return f->second(v);
            }
            throw QasmError(t.line, "unknown identifier '" + t.s + "' in an expression");
        }
        throw QasmError(t.line, "unexpected '" + t.s + "' in an expression");
    }
    std::vector<double> param_list(const Env& env)
    {
        std::vector<double> ps;
        if (!is_sym("(")) return ps;
        pos++;
        if (!is_sym(")")) {
            ps.push_back(expr(env));
            while (is_sym(",")) {
                pos++;
                ps.push_back(expr(env));
            }
        }
        expect_sym(")");
        return ps;
    }

    // ---- gate application ----
    void apply(const std::string& name, const std::vector<double>& ps,
               const std::vector<int>& qs, int line, int depth)
    {
        if (depth > 64) throw QasmError(line, "gate definitions nest too deeply");
        if (++applications > kMaxApplications)
            throw QasmError(line, "the program applies more than " +
                                      std::to_string(kMaxApplications) +
                                      " gates, user-defined ones included");
        auto d = defs.find(name);
        if (d != defs.end()) {
            const GateDef& g = d->second;
            if (g.opaque) throw QasmError(line, "opaque gate '" + name + "' has no definition");
            if (ps.size() != g.params.size() || qs.size() != g.qargs.size())
                throw QasmError(line, "gate '" + name + "' takes " +
                                          std::to_string(g.params.size()) + " parameters and " +
                                          std::to_string(g.qargs.size()) + " qubits");
            Env env;
            for (size_t i = 0; i < ps.size(); i++) env[g.params[i]] = ps[i];
            std::map<std::string, int> qa;
            for (size_t i = 0; i < qs.size(); i++) qa[g.qargs[i]] = qs[i];
            const size_t saved = pos;
            pos = g.body_begin;
            while (pos < g.body_end) body_statement(env, qa, depth + 1);
            pos = saved;
            return;
        }
        auto b = builtins().find(name);
        if (b == builtins().end()) throw QasmError(line, "unknown gate '" + name + "'");
        const int k = op_arity(b->second.op);
        if ((int)qs.size() != k || ps.size() != b->second.slots.size())
            throw QasmError(line, "gate '" + name + "' takes " +
                                      std::to_string(b->second.slots.size()) +
                                      " parameters and " + std::to_string(k) + " qubits");
        for (int i = 0; i < k; i++)
            for (int j = 0; j < i; j++)
                if (qs[i] == qs[j])
                    throw QasmError(line, "gate '" + name + "' applied to a repeated qubit");
        Gate g;
        g.op = b->second.op;
        for (int i = 0; i < k; i++) g.qb[i] = (uint32_t)qs[i];
        double slot[3] = {0, 0, 0};
        for (size_t i = 0; i < ps.size(); i++) slot[b->second.slots[i]] = ps[i];
        g.theta = slot[0];
        g.phi = slot[1];
        g.lambda = slot[2];
        if (prog->gates.size() >= kMaxGates)
            throw QasmError(line, "the program expands to more than " +
                                      std::to_string(kMaxGates) + " gates");
        prog->gates.push_back(std::move(g));
    }

    // statement inside a gate body: qubit names are the gate's arguments
    void body_statement(const Env& env, const std::map<std::string, int>& qa, int depth)
    {
        const int line = peek().line;
        const std::string name = expect_id();
        if (name == "barrier") {
            skip_statement();
            return;
        }
        const std::vector<double> ps = param_list(env);
        std::vector<int> qs;
        while (true) {
            const std::string a = expect_id();
            auto it = qa.find(a);
            if (it == qa.end()) throw QasmError(line, "'" + a + "' is not an argument of the gate");
            qs.push_back(it->second);
            if (!is_sym(",")) break;
            pos++;
        }
        expect_sym(";");
        apply(name, ps, qs, line, depth);
    }

    // top-level argument: reg[i] (one qubit) or reg (every qubit of reg)
    std::vector<int> qarg()
    {
        const int line = peek().line;
        const std::string r = expect_id();
        auto it = qregs.find(r);
        if (it == qregs.end()) throw QasmError(line, "unknown quantum register '" + r + "'");
        if (is_sym("[")) {
            pos++;
            const int i = expect_int();
            expect_sym("]");
            if (i >= it->second.second)
                throw QasmError(line, r + "[" + std::to_string(i) + "] is out of range");
            return {it->second.first + i};
        }
        std::vector<int> all;
        for (int i = 0; i < it->second.second; i++) all.push_back(it->second.first + i);
        return all;
    }

    void gate_def(bool opaque)
    {
        const int line = peek().line;
        const std::string name = expect_id();
        GateDef g;
        g.opaque = opaque;
        if (is_sym("(")) {
            pos++;
            if (!is_sym(")")) {
                g.params.push_back(expect_id());
                while (is_sym(",")) {
                    pos++;
                    g.params.push_back(expect_id());
                }
            }
            expect_sym(")");
        }
        g.qargs.push_back(expect_id());
        while (is_sym(",")) {
            pos++;
            g.qargs.push_back(expect_id());
        }
        if (opaque) {
            expect_sym(";");
        } else {
            expect_sym("{");
            g.body_begin = pos;
            int nest = 1;
            while (peek().k != T_END) {
                if (is_sym("{")) nest++;
                if (is_sym("}") && --nest == 0) break;
                pos++;
            }
            if (peek().k == T_END) throw QasmError(line, "unterminated body of gate '" + name + "'");
            g.body_end = pos;
            pos++;  // '}'
        }
        defs[name] = std::move(g);
    }

    void statement()
    {
        const int line = peek().line;
        if (peek().k != T_ID) throw QasmError(line, "unexpected '" + peek().s + "'");
        const std::string kw = peek().s;
        if (kw == "OPENQASM" || kw == "include" || kw == "measure" || kw == "barrier" ||
            kw == "reset" || kw == "if") {
            // the reference translator ignores these (tool/dmsim_qasm.py:105)
            skip_statement();
            return;
        }
        if (kw == "qreg" || kw == "creg") {
            pos++;
            const std::string r = expect_id();
            expect_sym("[");
            const int sz = expect_int();
            expect_sym("]");
            expect_sym(";");
            if (sz < 1) throw QasmError(line, "register '" + r + "' needs at least one bit");
            if (kw == "qreg") {
                if (qregs.count(r)) throw QasmError(line, "register '" + r + "' declared twice");
                qregs[r] = {prog->n_qubits, sz};
                prog->n_qubits += sz;
                if (prog->n_qubits > 64) throw QasmError(line, "more than 64 qubits");
            } else {
                cregs[r] = sz;
            }
            return;
        }
        if (kw == "gate" || kw == "opaque") {
            pos++;
            gate_def(kw == "opaque");
            return;
        }
        // gate application, broadcast over whole-register arguments
        pos++;
        const std::vector<double> ps = param_list(Env{});
        std::vector<std::vector<int>> args;
        args.push_back(qarg());
        while (is_sym(",")) {
            pos++;
            args.push_back(qarg());
        }
        expect_sym(";");
        size_t reps = 1;
        for (const auto& a : args)
            if (a.size() > 1) {
                if (reps > 1 && a.size() != reps)
                    throw QasmError(line, "registers of different sizes in one statement");
                reps = a.size();
            }
        for (size_t r = 0; r < reps; r++) {
            std::vector<int> qs;
            for (const auto& a : args) qs.push_back(a.size() > 1 ? a[r] : a[0]);
            apply(kw, ps, qs, line, 0);
        }
    }
};

}  // namespace

int parse_qasm(const std::string& src, QasmProgram& out)
{
    out = QasmProgram{};
    try {
        Parser p;
        p.toks = lex(src);
        p.prog = &out;
        while (p.peek().k != T_END) p.statement();
    } catch (const QasmError& e) {
        set_error(std::string("OpenQASM: ") + e.what());
        out = QasmProgram{};
        return DMX_EINVAL;
    }
    return DMX_OK;
}

}  // namespace dmx
