#include "expr.h"

#include <cctype>
#include <cmath>
#include <cstdlib>

namespace nmk {

struct Expr::Node {
    enum Kind { Num, Var, Neg, Bin, Fun, Tern } kind = Num;
    double value = 0;
    int var = -1;
    std::string op;  // binary operator or function name
    std::vector<std::unique_ptr<Node>> kids;

    double eval(const double *v) const {
        switch (kind) {
            case Num: return value;
            case Var: return v[var];
            case Neg: return -kids[0]->eval(v);
            case Tern: return kids[0]->eval(v) != 0 ? kids[1]->eval(v) : kids[2]->eval(v);
            case Bin: {
                if (op == "&&") return (kids[0]->eval(v) != 0 && kids[1]->eval(v) != 0) ? 1 : 0;
                if (op == "||") return (kids[0]->eval(v) != 0 || kids[1]->eval(v) != 0) ? 1 : 0;
                double a = kids[0]->eval(v), b = kids[1]->eval(v);
                if (op == "+") return a + b;
                if (op == "-") return a - b;
                if (op == "*") return a * b;
                if (op == "/") return a / b;
                if (op == "^") return kids[1]->kind == Num && kids[1]->value == 2.0 ? a * a : std::pow(a, b);
                if (op == "<") return a < b;
                if (op == ">") return a > b;
                if (op == "<=") return a <= b;
                if (op == ">=") return a >= b;
                if (op == "==") return a == b;
                if (op == "!=") return a != b;
                return 0;
            }
            case Fun: {
                std::vector<double> a;
                for (const auto &k : kids) a.push_back(k->eval(v));
                const std::string &f = op;
                if (f == "sin") return std::sin(a[0]);
                if (f == "cos") return std::cos(a[0]);
                if (f == "tan") return std::tan(a[0]);
                if (f == "asin") return std::asin(a[0]);
                if (f == "acos") return std::acos(a[0]);
                if (f == "atan") return std::atan(a[0]);
                if (f == "sinh") return std::sinh(a[0]);
                if (f == "cosh") return std::cosh(a[0]);
                if (f == "tanh") return std::tanh(a[0]);
                if (f == "asinh") return std::asinh(a[0]);
                if (f == "acosh") return std::acosh(a[0]);
                if (f == "atanh") return std::atanh(a[0]);
                if (f == "log2") return std::log2(a[0]);
                if (f == "log10") return std::log10(a[0]);
                if (f == "log" || f == "ln") return std::log(a[0]);
                if (f == "exp") return std::exp(a[0]);
                if (f == "sqrt") return std::sqrt(a[0]);
                if (f == "abs") return std::fabs(a[0]);
                if (f == "sign") return a[0] > 0 ? 1.0 : (a[0] < 0 ? -1.0 : 0.0);
                if (f == "rint") return std::rint(a[0]);
                if (f == "atan2") return std::atan2(a[0], a[1]);
                if (f == "sum" || f == "avg") {
                    double s = 0;
                    for (double x : a) s += x;
                    return f == "sum" ? s : s / double(a.size());
                }
                if (f == "min") {
                    double m = a[0];
                    for (double x : a) m = x < m ? x : m;
                    return m;
                }
                if (f == "max") {
                    double m = a[0];
                    for (double x : a) m = x > m ? x : m;
                    return m;
                }
                return 0;
            }
        }
        return 0;
    }
};

namespace {

struct FunSpec {
    const char *name;
    int min_args, max_args;  // max_args < 0: variadic
};
const FunSpec kFuns[] = {
    {"sin", 1, 1}, {"cos", 1, 1}, {"tan", 1, 1}, {"asin", 1, 1}, {"acos", 1, 1}, {"atan", 1, 1},
    {"sinh", 1, 1}, {"cosh", 1, 1}, {"tanh", 1, 1}, {"asinh", 1, 1}, {"acosh", 1, 1}, {"atanh", 1, 1},
    {"log2", 1, 1}, {"log10", 1, 1}, {"log", 1, 1}, {"ln", 1, 1}, {"exp", 1, 1}, {"sqrt", 1, 1},
    {"abs", 1, 1}, {"sign", 1, 1}, {"rint", 1, 1}, {"atan2", 2, 2},
    {"sum", 1, -1}, {"avg", 1, -1}, {"min", 1, -1}, {"max", 1, -1},
};

using NodeP = std::unique_ptr<Expr::Node>;

class Parser {
public:
    Parser(const std::string &s, const std::vector<std::string> &vars) : s_(s), vars_(vars) {}

    NodeP parse() {
        NodeP n = ternary();
        skip();
        if (pos_ != s_.size()) fail("unexpected token");
        return n;
    }

private:
    const std::string &s_;
    const std::vector<std::string> &vars_;
    size_t pos_ = 0;
    int depth_ = 0;  // nested sub-expressions and unary signs (recursion depth)

    // bounds the recursion: a formula nested deeper is refused rather than
    // overflowing the stack
    struct Nest {
        Parser &p;
        explicit Nest(Parser &q) : p(q) {
            if (++p.depth_ > 1000) p.fail("expression nested too deeply");
        }
        ~Nest() { --p.depth_; }
    };

    [[noreturn]] void fail(const std::string &m) const {
        throw ExprError("expression error at position " + std::to_string(pos_) + ": " + m + " in \"" + s_ + "\"");
    }
    void skip() {
        while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_])) pos_++;
    }
    bool eat(const char *tok) {
        skip();
        size_t n = std::string(tok).size();
        if (s_.compare(pos_, n, tok) == 0) {
            pos_ += n;
            return true;
        }
        return false;
    }
    static NodeP bin(const std::string &op, NodeP a, NodeP b) {
        NodeP n(new Expr::Node);
        n->kind = Expr::Node::Bin;
        n->op = op;
        n->kids.push_back(std::move(a));
        n->kids.push_back(std::move(b));
        return n;
    }

    NodeP ternary() {
        Nest guard(*this);
        NodeP c = lor();
        if (eat("?")) {
            NodeP a = ternary();
            if (!eat(":")) fail("expected ':'");
            NodeP b = ternary();
            NodeP n(new Expr::Node);
            n->kind = Expr::Node::Tern;
            n->kids.push_back(std::move(c));
            n->kids.push_back(std::move(a));
            n->kids.push_back(std::move(b));
            return n;
        }
        return c;
    }
    NodeP lor() {
        NodeP a = land();
        while (eat("||")) a = bin("||", std::move(a), land());
        return a;
    }
    NodeP land() {
        NodeP a = cmp();
        while (eat("&&")) a = bin("&&", std::move(a), cmp());
        return a;
    }
    NodeP cmp() {
        NodeP a = add();
        for (;;) {
            const char *ops[] = {"<=", ">=", "==", "!=", "<", ">"};
            bool hit = false;
            for (const char *o : ops)
                if (eat(o)) {
                    a = bin(o, std::move(a), add());
                    hit = true;
                    break;
                }
            if (!hit) return a;
        }
    }
    NodeP add() {
        NodeP a = mul();
        for (;;) {
            if (eat("+")) a = bin("+", std::move(a), mul());
            else if (eat("-")) a = bin("-", std::move(a), mul());
            else return a;
        }
    }
    NodeP mul() {
        NodeP a = unary();
        for (;;) {
            if (eat("*")) a = bin("*", std::move(a), unary());
            else if (eat("/")) a = bin("/", std::move(a), unary());
            else return a;
        }
    }
    // unary sign binds looser than ^ : -2^2 = -4
    NodeP unary() {
        Nest guard(*this);
        if (eat("-")) {
            NodeP n(new Expr::Node);
            n->kind = Expr::Node::Neg;
            n->kids.push_back(unary());
            return n;
        }
        if (eat("+")) return unary();
        return power();
    }
    NodeP power() {
        NodeP base = primary();
        if (eat("^")) return bin("^", std::move(base), unary());  // right associative
        return base;
    }
    NodeP primary() {
        skip();
        if (pos_ >= s_.size()) fail("unexpected end");
        char c = s_[pos_];
        if (c == '(') {
            pos_++;
            NodeP n = ternary();
            if (!eat(")")) fail("expected ')'");
            return n;
        }
        if (std::isdigit((unsigned char)c) || c == '.') {
            const char *begin = s_.c_str() + pos_;
            char *end = nullptr;
            double v = std::strtod(begin, &end);
            if (end == begin) fail("invalid number");
            pos_ += size_t(end - begin);
            NodeP n(new Expr::Node);
            n->value = v;
            return n;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t st = pos_;
            while (pos_ < s_.size() && (std::isalnum((unsigned char)s_[pos_]) || s_[pos_] == '_')) pos_++;
            std::string name = s_.substr(st, pos_ - st);
            skip();
            if (pos_ < s_.size() && s_[pos_] == '(') {
                const FunSpec *spec = nullptr;
                for (const auto &f : kFuns)
                    if (name == f.name) spec = &f;
                if (!spec) fail("unknown function '" + name + "'");
                pos_++;
                NodeP n(new Expr::Node);
                n->kind = Expr::Node::Fun;
                n->op = name;
                if (!eat(")")) {
                    for (;;) {
                        n->kids.push_back(ternary());
                        if (eat(",")) continue;
                        if (!eat(")")) fail("expected ')'");
                        break;
                    }
                }
                int na = int(n->kids.size());
                if (na < spec->min_args || (spec->max_args >= 0 && na > spec->max_args))
                    fail("wrong number of arguments for '" + name + "'");
                return n;
            }
            for (size_t i = 0; i < vars_.size(); i++)
                if (vars_[i] == name) {
                    NodeP n(new Expr::Node);
                    n->kind = Expr::Node::Var;
                    n->var = int(i);
                    return n;
                }
            if (name == "_pi") {
                NodeP n(new Expr::Node);
                n->value = 3.141592653589793238462643;
                return n;
            }
            if (name == "_e") {
                NodeP n(new Expr::Node);
                n->value = 2.718281828459045235360287;
                return n;
            }
            fail("unknown variable '" + name + "'");
        }
        fail(std::string("unexpected character '") + c + "'");
    }
};

}  // namespace

Expr::Expr(const std::string &text, const std::vector<std::string> &variables) {
    Parser p(text, variables);
    root_ = p.parse();
}
Expr::~Expr() = default;
Expr::Expr(Expr &&) noexcept = default;
Expr &Expr::operator=(Expr &&) noexcept = default;

double Expr::eval(const double *values) const { return root_->eval(values); }

namespace {

int fun1_code(const std::string &f) {
    static const char *names[] = {"sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh", "tanh", "asinh",
                                  "acosh", "atanh", "log2", "log10", "log", "exp", "sqrt", "abs", "sign", "rint"};
    for (int i = 0; i < int(sizeof(names) / sizeof(names[0])); i++)
        if (f == names[i]) return i;
    if (f == "ln") return EF_LOG;
    return -1;
}

int bin_code(const std::string &op) {
    static const struct {
        const char *op;
        int code;
    } t[] = {{"+", EX_ADD}, {"-", EX_SUB}, {"*", EX_MUL}, {"/", EX_DIV}, {"^", EX_POW}, {"<", EX_LT},
             {">", EX_GT}, {"<=", EX_LE}, {">=", EX_GE}, {"==", EX_EQ}, {"!=", EX_NE}, {"&&", EX_AND},
             {"||", EX_OR}};
    for (const auto &e : t)
        if (op == e.op) return e.code;
    return -1;
}

// appends the postfix form of n; depth: stack slots in use before n
bool emit(const Expr::Node &n, ExprProgram &p, int depth) {
    auto push = [&](int code, int arg, double value, int new_depth) {
        if (p.n >= kExprMaxOps || new_depth > kExprMaxStack) return false;
        p.ops[p.n++] = ExprOp{code, arg, value};
        if (new_depth > p.depth) p.depth = new_depth;
        return true;
    };
    using K = Expr::Node;
    switch (n.kind) {
        case K::Num: return push(EX_CONST, 0, n.value, depth + 1);
        case K::Var: return push(EX_VAR, n.var, 0, depth + 1);
        case K::Neg: return emit(*n.kids[0], p, depth) && push(EX_NEG, 0, 0, depth + 1);
        case K::Tern:
            return emit(*n.kids[0], p, depth) && emit(*n.kids[1], p, depth + 1) && emit(*n.kids[2], p, depth + 2) &&
                   push(EX_TERN, 0, 0, depth + 1);
        case K::Bin: {
            const int code = bin_code(n.op);
            if (code < 0) return false;
            if (code == EX_POW && n.kids[1]->kind == K::Num && n.kids[1]->value == 2.0)
                return emit(*n.kids[0], p, depth) && push(EX_SQR, 0, 0, depth + 1);
            return emit(*n.kids[0], p, depth) && emit(*n.kids[1], p, depth + 1) && push(code, 0, 0, depth + 1);
        }
        case K::Fun: {
            const int na = int(n.kids.size());
            for (int i = 0; i < na; i++)
                if (!emit(*n.kids[i], p, depth + i)) return false;
            if (n.op == "atan2") return push(EX_ATAN2, 0, 0, depth + 1);
            if (n.op == "sum") return push(EX_SUM, na, 0, depth + 1);
            if (n.op == "avg") return push(EX_AVG, na, 0, depth + 1);
            if (n.op == "min") return push(EX_MIN, na, 0, depth + 1);
            if (n.op == "max") return push(EX_MAX, na, 0, depth + 1);
            const int f = fun1_code(n.op);
            return f >= 0 && na == 1 && push(EX_FUN1, f, 0, depth + 1);
        }
    }
    return false;
}

}  // namespace

bool Expr::compile(ExprProgram &out) const {
    ExprProgram p;
    if (!emit(*root_, p, 0)) return false;
    out = p;
    return true;
}

double eval_index_expression(const std::string &text, double lambda) {
    Expr e(text, {"lambda"});
    return e.eval(&lambda);
}

}  // namespace nmk
