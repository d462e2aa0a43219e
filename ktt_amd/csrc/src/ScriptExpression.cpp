#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <string>

#include <KttException.h>

#include "ScriptExpression.h"

namespace ktt
{

// ---------------------------------------------------------------- values

ScriptValue ScriptValue::Integer(const int64_t value)
{
    ScriptValue result;
    result.m_Kind = Kind::Int;
    result.m_Int = value;
    return result;
}

ScriptValue ScriptValue::Real(const double value)
{
    ScriptValue result;
    result.m_Kind = Kind::Float;
    result.m_Float = value;
    return result;
}

ScriptValue ScriptValue::Boolean(const bool value)
{
    ScriptValue result;
    result.m_Kind = Kind::Bool;
    result.m_Int = value ? 1 : 0;
    return result;
}

ScriptValue ScriptValue::Sequence(std::vector<ScriptValue> items)
{
    ScriptValue result;
    result.m_Kind = Kind::List;
    result.m_Items = std::move(items);
    return result;
}

double ScriptValue::ToDouble() const
{
    switch (m_Kind)
    {
    case Kind::Float:
        return m_Float;
    case Kind::List:
        throw KttException("A list is used where a number is expected");
    default:
        return static_cast<double>(m_Int);
    }
}

int64_t ScriptValue::ToInt() const
{
    if (m_Kind == Kind::List)
    {
        throw KttException("A list is used where an integer is expected");
    }
    if (m_Kind == Kind::Float)
    {
        if (std::floor(m_Float) != m_Float)
        {
            throw KttException("The value " + ToString() + " is not an integer");
        }
        return static_cast<int64_t>(m_Float);
    }
    return m_Int;
}

bool ScriptValue::Truthy() const
{
    switch (m_Kind)
    {
    case Kind::Float:
        return m_Float != 0.0;
    case Kind::List:
        return !m_Items.empty();
    default:
        return m_Int != 0;
    }
}

const std::vector<ScriptValue>& ScriptValue::Items() const
{
    if (m_Kind != Kind::List)
    {
        throw KttException("The value " + ToString() + " is not a list");
    }
    return m_Items;
}

std::string ScriptValue::ToString() const
{
    switch (m_Kind)
    {
    case Kind::Int:
        return std::to_string(m_Int);
    case Kind::Float:
        return std::to_string(m_Float);
    case Kind::Bool:
        return m_Int != 0 ? "True" : "False";
    default:
    {
        std::string text = "[";
        for (size_t i = 0; i < m_Items.size(); ++i)
        {
            text += (i == 0 ? "" : ", ") + m_Items[i].ToString();
        }
        return text + "]";
    }
    }
}

// ---------------------------------------------------------------- syntax tree

struct ScriptExpression::Node
{
    enum class Type
    {
        Literal,
        Name,
        Unary,
        Binary,
        Compare,
        And,
        Or,
        Not,
        Conditional,
        List,
        Call
    };

    Type type;
    std::string text;                          // operator, name or function
    ScriptValue value;                         // literal
    std::vector<std::unique_ptr<Node>> children;
    std::vector<std::string> comparisons;      // operators of a comparison chain
};

namespace
{

using Node = ScriptExpression::Node;

struct Token
{
    enum class Type
    {
        Number,
        Name,
        Operator,
        End
    };

    Type type;
    std::string text;
    size_t position;
};

std::vector<Token> Tokenize(const std::string& text)
{
    static const std::vector<std::string> operators = {"**", "//", "<=", ">=", "==", "!=", "+", "-", "*", "/", "%", "<", ">", "(",
        ")", "[", "]", ","};
    std::vector<Token> tokens;
    size_t at = 0;
    while (at < text.size())
    {
        const char c = text[at];
        if (std::isspace(static_cast<unsigned char>(c)))
        {
            ++at;
            continue;
        }
        if (std::isdigit(static_cast<unsigned char>(c)) || (c == '.' && at + 1 < text.size() && std::isdigit(static_cast<unsigned char>(text[at + 1]))))
        {
            size_t end = at;
            while (end < text.size() && (std::isdigit(static_cast<unsigned char>(text[end])) || text[end] == '.'))
            {
                ++end;
            }
            if (end < text.size() && (text[end] == 'e' || text[end] == 'E'))
            {
                size_t exponent = end + 1;
                if (exponent < text.size() && (text[exponent] == '+' || text[exponent] == '-'))
                {
                    ++exponent;
                }
                if (exponent < text.size() && std::isdigit(static_cast<unsigned char>(text[exponent])))
                {
                    end = exponent;
                    while (end < text.size() && std::isdigit(static_cast<unsigned char>(text[end])))
                    {
                        ++end;
                    }
                }
            }
            tokens.push_back({Token::Type::Number, text.substr(at, end - at), at});
            at = end;
            continue;
        }
        if (std::isalpha(static_cast<unsigned char>(c)) || c == '_')
        {
            size_t end = at;
            while (end < text.size() && (std::isalnum(static_cast<unsigned char>(text[end])) || text[end] == '_'))
            {
                ++end;
            }
            tokens.push_back({Token::Type::Name, text.substr(at, end - at), at});
            at = end;
            continue;
        }
        const auto match = std::find_if(operators.begin(), operators.end(),
            [&](const std::string& op) { return text.compare(at, op.size(), op) == 0; });
        if (match == operators.end())
        {
            throw KttException("Unexpected character '" + std::string(1, c) + "' at " + std::to_string(at) + " in \"" + text + "\"");
        }
        tokens.push_back({Token::Type::Operator, *match, at});
        at += match->size();
    }
    tokens.push_back({Token::Type::End, "", text.size()});
    return tokens;
}

std::unique_ptr<Node> MakeNode(const Node::Type type, const std::string& text = "")
{
    auto node = std::make_unique<Node>();
    node->type = type;
    node->text = text;
    return node;
}

/** Recursive descent over Python's expression grammar, lowest precedence first. */
class Parser
{
public:
    Parser(const std::string& text, std::vector<std::string>& names) :
        m_Text(text),
        m_Tokens(Tokenize(text)),
        m_Next(0),
        m_Names(names)
    {}

    std::unique_ptr<Node> ParseAll()
    {
        auto root = Conditional();
        if (Peek().type != Token::Type::End)
        {
            Fail("unexpected '" + Peek().text + "'");
        }
        return root;
    }

private:
    const std::string& m_Text;
    std::vector<Token> m_Tokens;
    size_t m_Next;
    std::vector<std::string>& m_Names;

    const Token& Peek() const { return m_Tokens[m_Next]; }
    bool IsOperator(const std::string& op) const { return Peek().type == Token::Type::Operator && Peek().text == op; }
    bool IsKeyword(const std::string& word) const { return Peek().type == Token::Type::Name && Peek().text == word; }

    [[noreturn]] void Fail(const std::string& what) const
    {
        throw KttException("Syntax error at " + std::to_string(Peek().position) + " in \"" + m_Text + "\": " + what);
    }

    void Expect(const std::string& op)
    {
        if (!IsOperator(op))
        {
            Fail("expected '" + op + "'");
        }
        ++m_Next;
    }

    std::unique_ptr<Node> Conditional()
    {
        auto value = Or();
        if (!IsKeyword("if"))
        {
            return value;
        }
        ++m_Next;
        auto condition = Or();
        if (!IsKeyword("else"))
        {
            Fail("expected 'else'");
        }
        ++m_Next;
        auto otherwise = Conditional();
        auto node = MakeNode(Node::Type::Conditional);
        node->children.push_back(std::move(condition));
        node->children.push_back(std::move(value));
        node->children.push_back(std::move(otherwise));
        return node;
    }

    std::unique_ptr<Node> Or()
    {
        auto left = And();
        while (IsKeyword("or"))
        {
            ++m_Next;
            auto node = MakeNode(Node::Type::Or);
            node->children.push_back(std::move(left));
            node->children.push_back(And());
            left = std::move(node);
        }
        return left;
    }

    std::unique_ptr<Node> And()
    {
        auto left = Not();
        while (IsKeyword("and"))
        {
            ++m_Next;
            auto node = MakeNode(Node::Type::And);
            node->children.push_back(std::move(left));
            node->children.push_back(Not());
            left = std::move(node);
        }
        return left;
    }

    std::unique_ptr<Node> Not()
    {
        if (IsKeyword("not"))
        {
            ++m_Next;
            auto node = MakeNode(Node::Type::Not);
            node->children.push_back(Not());
            return node;
        }
        return Comparison();
    }

    std::unique_ptr<Node> Comparison()
    {
        auto first = Sum();
        static const std::vector<std::string> relations = {"<", "<=", ">", ">=", "==", "!="};
        if (Peek().type != Token::Type::Operator || std::find(relations.begin(), relations.end(), Peek().text) == relations.end())
        {
            return first;
        }
        auto node = MakeNode(Node::Type::Compare);
        node->children.push_back(std::move(first));
        while (Peek().type == Token::Type::Operator && std::find(relations.begin(), relations.end(), Peek().text) != relations.end())
        {
            node->comparisons.push_back(Peek().text);
            ++m_Next;
            node->children.push_back(Sum());
        }
        return node;
    }

    std::unique_ptr<Node> Sum()
    {
        auto left = Product();
        while (IsOperator("+") || IsOperator("-"))
        {
            auto node = MakeNode(Node::Type::Binary, Peek().text);
            ++m_Next;
            node->children.push_back(std::move(left));
            node->children.push_back(Product());
            left = std::move(node);
        }
        return left;
    }

    std::unique_ptr<Node> Product()
    {
        auto left = Unary();
        while (IsOperator("*") || IsOperator("/") || IsOperator("//") || IsOperator("%"))
        {
            auto node = MakeNode(Node::Type::Binary, Peek().text);
            ++m_Next;
            node->children.push_back(std::move(left));
            node->children.push_back(Unary());
            left = std::move(node);
        }
        return left;
    }

    std::unique_ptr<Node> Unary()
    {
        if (IsOperator("-") || IsOperator("+"))
        {
            auto node = MakeNode(Node::Type::Unary, Peek().text);
            ++m_Next;
            node->children.push_back(Unary());
            return node;
        }
        return Power();
    }

    std::unique_ptr<Node> Power()
    {
        auto base = Atom();
        if (!IsOperator("**"))
        {
            return base;
        }
        ++m_Next;
        auto node = MakeNode(Node::Type::Binary, "**");
        node->children.push_back(std::move(base));
        node->children.push_back(Unary());  // right-associative, binds tighter than unary on its left
        return node;
    }

    std::unique_ptr<Node> Atom()
    {
        const Token token = Peek();
        if (token.type == Token::Type::Number)
        {
            ++m_Next;
            auto node = MakeNode(Node::Type::Literal);
            const bool real = token.text.find_first_of(".eE") != std::string::npos;
            node->value = real ? ScriptValue::Real(std::strtod(token.text.c_str(), nullptr))
                               : ScriptValue::Integer(std::strtoll(token.text.c_str(), nullptr, 10));
            return node;
        }
        if (token.type == Token::Type::Name)
        {
            ++m_Next;
            if (token.text == "True" || token.text == "False")
            {
                auto node = MakeNode(Node::Type::Literal);
                node->value = ScriptValue::Boolean(token.text == "True");
                return node;
            }
            if (IsOperator("("))
            {
                ++m_Next;
                auto node = MakeNode(Node::Type::Call, token.text);
                if (!IsOperator(")"))
                {
                    node->children.push_back(Conditional());
                    while (IsOperator(","))
                    {
                        ++m_Next;
                        node->children.push_back(Conditional());
                    }
                }
                Expect(")");
                return node;
            }
            if (std::find(m_Names.begin(), m_Names.end(), token.text) == m_Names.end())
            {
                m_Names.push_back(token.text);
            }
            return MakeNode(Node::Type::Name, token.text);
        }
        if (IsOperator("("))
        {
            ++m_Next;
            auto inner = Conditional();
            Expect(")");
            return inner;
        }
        if (IsOperator("["))
        {
            ++m_Next;
            auto node = MakeNode(Node::Type::List);
            if (!IsOperator("]"))
            {
                node->children.push_back(Conditional());
                while (IsOperator(","))
                {
                    ++m_Next;
                    if (IsOperator("]"))
                    {
                        break;  // trailing comma
                    }
                    node->children.push_back(Conditional());
                }
            }
            Expect("]");
            return node;
        }
        Fail(token.type == Token::Type::End ? "unexpected end" : "unexpected '" + token.text + "'");
    }
};

bool BothIntegral(const ScriptValue& a, const ScriptValue& b)
{
    return a.GetKind() != ScriptValue::Kind::Float && b.GetKind() != ScriptValue::Kind::Float;
}

// Python integers are unbounded; this evaluator keeps int64 and refuses a result outside it
// rather than wrapping.
int64_t CheckedAdd(const int64_t a, const int64_t b)
{
    int64_t result = 0;
    if (__builtin_add_overflow(a, b, &result))
    {
        throw KttException("Integer overflow in " + std::to_string(a) + " + " + std::to_string(b));
    }
    return result;
}

int64_t CheckedSubtract(const int64_t a, const int64_t b)
{
    int64_t result = 0;
    if (__builtin_sub_overflow(a, b, &result))
    {
        throw KttException("Integer overflow in " + std::to_string(a) + " - " + std::to_string(b));
    }
    return result;
}

int64_t CheckedMultiply(const int64_t a, const int64_t b)
{
    int64_t result = 0;
    if (__builtin_mul_overflow(a, b, &result))
    {
        throw KttException("Integer overflow in " + std::to_string(a) + " * " + std::to_string(b));
    }
    return result;
}

// base ** exponent for exponent >= 0 by square-and-multiply, every step checked.
int64_t CheckedPower(int64_t base, int64_t exponent)
{
    int64_t result = 1;
    while (exponent > 0)
    {
        if (exponent & 1)
        {
            result = CheckedMultiply(result, base);
        }
        exponent >>= 1;
        if (exponent > 0)
        {
            base = CheckedMultiply(base, base);
        }
    }
    return result;
}

int64_t FloorDivide(const int64_t a, const int64_t b)
{
    if (b == 0)
    {
        throw KttException("Integer division by zero");
    }
    if (a == INT64_MIN && b == -1)
    {
        throw KttException("Integer overflow in " + std::to_string(a) + " // -1");
    }
    int64_t quotient = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0)))
    {
        --quotient;
    }
    return quotient;
}

/** Python's float floor division and modulo (the remainder has the divisor's sign). */
void FloatDivMod(const double a, const double b, double& floorQuotient, double& remainder)
{
    if (b == 0.0)
    {
        throw KttException("Float division by zero");
    }
    double mod = std::fmod(a, b);
    double div = (a - mod) / b;
    if (mod != 0.0)
    {
        if ((b < 0.0) != (mod < 0.0))
        {
            mod += b;
            div -= 1.0;
        }
    }
    else
    {
        mod = std::copysign(0.0, b);
    }
    if (div != 0.0)
    {
        floorQuotient = std::floor(div);
        if (div - floorQuotient > 0.5)
        {
            floorQuotient += 1.0;
        }
    }
    else
    {
        floorQuotient = std::copysign(0.0, a / b);
    }
    remainder = mod;
}

ScriptValue Arithmetic(const std::string& op, const ScriptValue& a, const ScriptValue& b)
{
    if (!a.IsNumber() || !b.IsNumber())
    {
        if (op == "+" && a.GetKind() == ScriptValue::Kind::List && b.GetKind() == ScriptValue::Kind::List)
        {
            std::vector<ScriptValue> joined = a.Items();
            joined.insert(joined.end(), b.Items().begin(), b.Items().end());
            return ScriptValue::Sequence(std::move(joined));
        }
        throw KttException("Operator " + op + " needs numbers, got " + a.ToString() + " and " + b.ToString());
    }
    const bool integral = BothIntegral(a, b);
    if (op == "/")
    {
        if (b.ToDouble() == 0.0)
        {
            throw KttException("Division by zero");
        }
        return ScriptValue::Real(a.ToDouble() / b.ToDouble());
    }
    if (op == "//" || op == "%")
    {
        if (integral)
        {
            const int64_t x = a.ToInt();
            const int64_t y = b.ToInt();
            if (op == "%" && y == -1)
            {
                return ScriptValue::Integer(0);
            }
            const int64_t q = FloorDivide(x, y);
            return ScriptValue::Integer(op == "//" ? q : x - q * y);
        }
        double quotient = 0.0;
        double remainder = 0.0;
        FloatDivMod(a.ToDouble(), b.ToDouble(), quotient, remainder);
        return ScriptValue::Real(op == "//" ? quotient : remainder);
    }
    if (op == "**")
    {
        if (integral && b.ToInt() >= 0)
        {
            return ScriptValue::Integer(CheckedPower(a.ToInt(), b.ToInt()));
        }
        return ScriptValue::Real(std::pow(a.ToDouble(), b.ToDouble()));
    }
    if (integral)
    {
        const int64_t x = a.ToInt();
        const int64_t y = b.ToInt();
        if (op == "+")
        {
            return ScriptValue::Integer(CheckedAdd(x, y));
        }
        if (op == "-")
        {
            return ScriptValue::Integer(CheckedSubtract(x, y));
        }
        return ScriptValue::Integer(CheckedMultiply(x, y));
    }
    const double x = a.ToDouble();
    const double y = b.ToDouble();
    if (op == "+")
    {
        return ScriptValue::Real(x + y);
    }
    if (op == "-")
    {
        return ScriptValue::Real(x - y);
    }
    return ScriptValue::Real(x * y);
}

bool Relation(const std::string& op, const ScriptValue& a, const ScriptValue& b)
{
    if (!a.IsNumber() || !b.IsNumber())
    {
        if (op == "==" || op == "!=")
        {
            const bool same = a.ToString() == b.ToString();
            return op == "==" ? same : !same;
        }
        throw KttException("Lists cannot be ordered");
    }
    if (BothIntegral(a, b))
    {
        const int64_t x = a.ToInt();
        const int64_t y = b.ToInt();
        return op == "<" ? x < y : op == "<=" ? x <= y : op == ">" ? x > y : op == ">=" ? x >= y : op == "==" ? x == y : x != y;
    }
    const double x = a.ToDouble();
    const double y = b.ToDouble();
    return op == "<" ? x < y : op == "<=" ? x <= y : op == ">" ? x > y : op == ">=" ? x >= y : op == "==" ? x == y : x != y;
}

ScriptValue Call(const std::string& function, const std::vector<ScriptValue>& arguments)
{
    const auto numbers = [&]() -> std::vector<ScriptValue>
    {
        if (arguments.size() == 1 && arguments[0].GetKind() == ScriptValue::Kind::List)
        {
            return arguments[0].Items();
        }
        return arguments;
    };
    if (function == "range")
    {
        if (arguments.empty() || arguments.size() > 3)
        {
            throw KttException("range takes 1 to 3 arguments");
        }
        const int64_t start = arguments.size() == 1 ? 0 : arguments[0].ToInt();
        const int64_t stop = arguments.size() == 1 ? arguments[0].ToInt() : arguments[1].ToInt();
        const int64_t step = arguments.size() == 3 ? arguments[2].ToInt() : 1;
        if (step == 0)
        {
            throw KttException("range step must not be zero");
        }
        std::vector<ScriptValue> items;
        for (int64_t k = start; step > 0 ? k < stop : k > stop; k += step)
        {
            items.push_back(ScriptValue::Integer(k));
        }
        return ScriptValue::Sequence(std::move(items));
    }
    if (function == "len" && arguments.size() == 1)
    {
        return ScriptValue::Integer(static_cast<int64_t>(arguments[0].Items().size()));
    }
    if ((function == "min" || function == "max") && !arguments.empty())
    {
        const std::vector<ScriptValue> values = numbers();
        if (values.empty())
        {
            throw KttException(function + " of an empty sequence");
        }
        ScriptValue best = values[0];
        for (const ScriptValue& value : values)
        {
            if (function == "min" ? Relation("<", value, best) : Relation(">", value, best))
            {
                best = value;
            }
        }
        return best;
    }
    if (function == "abs" && arguments.size() == 1)
    {
        const ScriptValue& value = arguments[0];
        return value.GetKind() == ScriptValue::Kind::Float ? ScriptValue::Real(std::fabs(value.ToDouble()))
                                                            : ScriptValue::Integer(std::llabs(value.ToInt()));
    }
    if (function == "int" && arguments.size() == 1)
    {
        return ScriptValue::Integer(static_cast<int64_t>(std::trunc(arguments[0].ToDouble())));
    }
    if (function == "float" && arguments.size() == 1)
    {
        return ScriptValue::Real(arguments[0].ToDouble());
    }
    throw KttException("Unknown function " + function + " with " + std::to_string(arguments.size()) + " arguments");
}

ScriptValue Evaluate(const Node& node, const ScriptExpression::Lookup& lookup)
{
    switch (node.type)
    {
    case Node::Type::Literal:
        return node.value;
    case Node::Type::Name:
        return lookup(node.text);
    case Node::Type::Unary:
    {
        const ScriptValue operand = Evaluate(*node.children[0], lookup);
        if (node.text == "+")
        {
            return operand.GetKind() == ScriptValue::Kind::Bool ? ScriptValue::Integer(operand.ToInt()) : operand;
        }
        return operand.GetKind() == ScriptValue::Kind::Float ? ScriptValue::Real(-operand.ToDouble())
                                                             : ScriptValue::Integer(CheckedSubtract(0, operand.ToInt()));
    }
    case Node::Type::Binary:
        return Arithmetic(node.text, Evaluate(*node.children[0], lookup), Evaluate(*node.children[1], lookup));
    case Node::Type::Compare:
    {
        ScriptValue left = Evaluate(*node.children[0], lookup);
        for (size_t k = 0; k < node.comparisons.size(); ++k)
        {
            ScriptValue right = Evaluate(*node.children[k + 1], lookup);
            if (!Relation(node.comparisons[k], left, right))
            {
                return ScriptValue::Boolean(false);
            }
            left = right;
        }
        return ScriptValue::Boolean(true);
    }
    case Node::Type::And:
    {
        ScriptValue left = Evaluate(*node.children[0], lookup);
        return left.Truthy() ? Evaluate(*node.children[1], lookup) : left;
    }
    case Node::Type::Or:
    {
        ScriptValue left = Evaluate(*node.children[0], lookup);
        return left.Truthy() ? left : Evaluate(*node.children[1], lookup);
    }
    case Node::Type::Not:
        return ScriptValue::Boolean(!Evaluate(*node.children[0], lookup).Truthy());
    case Node::Type::Conditional:
        return Evaluate(*node.children[0], lookup).Truthy() ? Evaluate(*node.children[1], lookup) : Evaluate(*node.children[2], lookup);
    case Node::Type::List:
    {
        std::vector<ScriptValue> items;
        for (const auto& child : node.children)
        {
            items.push_back(Evaluate(*child, lookup));
        }
        return ScriptValue::Sequence(std::move(items));
    }
    case Node::Type::Call:
    {
        std::vector<ScriptValue> arguments;
        for (const auto& child : node.children)
        {
            arguments.push_back(Evaluate(*child, lookup));
        }
        return Call(node.text, arguments);
    }
    }
    throw KttException("Unknown expression node");
}

} // namespace

ScriptExpression::ScriptExpression(const std::string& text) :
    m_Text(text)
{
    Parser parser(m_Text, m_Names);
    m_Root = parser.ParseAll();
}

ScriptExpression::~ScriptExpression() = default;
ScriptExpression::ScriptExpression(ScriptExpression&&) noexcept = default;
ScriptExpression& ScriptExpression::operator=(ScriptExpression&&) noexcept = default;

ScriptValue ScriptExpression::Evaluate(const Lookup& lookup) const
{
    return ktt::Evaluate(*m_Root, lookup);
}

} // namespace ktt
