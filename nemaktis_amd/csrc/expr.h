// Expression evaluator for the index strings ("no", "ne", ..., evaluated in
// `lambda`) and the mask formula (evaluated in x, y, z) of the bpm-solver
// settings.  The reference evaluates both with muParser
// (BeamPropagationSolver/src/PhysicsLib/PhysicsCoefficients.cpp:27-103,
// UtilitiesLib/VectorField.cpp:72-104); this is an independent evaluator of
// the muParser syntax the nemaktis caller produces: numbers, variables,
// the constants _pi and _e, + - * / ^ (right associative, binding tighter
// than unary minus), comparisons, && ||, the ternary ?:, and muParser's
// built-in functions.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "expr_program.h"

namespace nmk {

struct ExprError : std::runtime_error {
    explicit ExprError(const std::string &m) : std::runtime_error(m) {}
};

class Expr {
public:
    // variables: names that may appear; their values are passed to eval() in
    // the same order
    Expr(const std::string &text, const std::vector<std::string> &variables);
    ~Expr();
    Expr(Expr &&) noexcept;
    Expr &operator=(Expr &&) noexcept;

    double eval(const double *values) const;
    double eval() const { return eval(nullptr); }

    // the postfix program of the expression (expr_program.h); false when it
    // needs more than kExprMaxOps operations or kExprMaxStack stack slots
    bool compile(ExprProgram &out) const;

    struct Node;

private:
    std::unique_ptr<Node> root_;
};

// evaluates `text` with the single variable `lambda` (wavelength in um)
double eval_index_expression(const std::string &text, double lambda);

}  // namespace nmk
