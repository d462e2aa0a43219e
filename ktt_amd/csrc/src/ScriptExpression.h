// Expressions of KTT tuning scripts: parameter value lists ("[16, 32]"), constraints
// ("WORK_GROUP_SIZE_X * WORK_GROUP_SIZE_Y >= 64"), launch sizes ("256 // Z_ITERATIONS") and
// element generators ("1+i*0.1 if i%4 < 3 else (i%10)*0.1"). The reference evaluates them
// with an embedded Python interpreter; this is an evaluator of the Python expression subset
// those scripts use, with Python's semantics: integers and floats are distinct, // and % floor
// towards minus infinity (the remainder takes the divisor's sign), / always gives a float,
// comparisons chain, "and"/"or" return an operand, and "x if c else y" is the conditional.
// Supported: integer, float, True/False literals, names, + - * / // % ** (and unary + -),
// < <= > >= == !=, not/and/or, conditional expressions, parentheses, list literals and the
// functions range, len, min, max, abs, int, float.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace ktt
{

class ScriptValue
{
public:
    enum class Kind
    {
        Int,
        Float,
        Bool,
        List
    };

    ScriptValue() : m_Kind(Kind::Int), m_Int(0), m_Float(0.0) {}
    static ScriptValue Integer(const int64_t value);
    static ScriptValue Real(const double value);
    static ScriptValue Boolean(const bool value);
    static ScriptValue Sequence(std::vector<ScriptValue> items);

    Kind GetKind() const { return m_Kind; }
    bool IsNumber() const { return m_Kind != Kind::List; }
    /** The value as a double (bools are 0/1); throws for a list. */
    double ToDouble() const;
    /** The value as an integer; floats must be integral; throws for a list. */
    int64_t ToInt() const;
    bool Truthy() const;
    const std::vector<ScriptValue>& Items() const;
    std::string ToString() const;

private:
    Kind m_Kind;
    int64_t m_Int;
    double m_Float;
    std::vector<ScriptValue> m_Items;
};

class ScriptExpression
{
public:
    using Lookup = std::function<ScriptValue(const std::string& name)>;

    /** Parses the text; throws KttException with the position on a syntax error. */
    explicit ScriptExpression(const std::string& text);
    ~ScriptExpression();
    ScriptExpression(ScriptExpression&&) noexcept;
    ScriptExpression& operator=(ScriptExpression&&) noexcept;

    ScriptValue Evaluate(const Lookup& lookup) const;
    /** The free names the expression reads (function names excluded), each once. */
    const std::vector<std::string>& Names() const { return m_Names; }
    const std::string& Text() const { return m_Text; }

    struct Node;

private:
    std::string m_Text;
    std::unique_ptr<Node> m_Root;
    std::vector<std::string> m_Names;
};

} // namespace ktt
