#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

#include <KttException.h>

#include "Json.h"

namespace ktt
{

JsonValue::JsonValue() : m_Type(Type::Null), m_Bool(false), m_Number(0.0), m_Integral(false), m_Negative(false), m_Magnitude(0) {}
JsonValue::JsonValue(std::nullptr_t) : JsonValue() {}

JsonValue::JsonValue(bool value) : JsonValue()
{
    m_Type = Type::Bool;
    m_Bool = value;
}

JsonValue::JsonValue(int value) : JsonValue(static_cast<int64_t>(value)) {}

JsonValue::JsonValue(int64_t value) : JsonValue()
{
    m_Type = Type::Number;
    m_Number = static_cast<double>(value);
    m_Integral = true;
    m_Negative = value < 0;
    m_Magnitude = value < 0 ? static_cast<uint64_t>(-(value + 1)) + 1 : static_cast<uint64_t>(value);
}

JsonValue::JsonValue(uint64_t value) : JsonValue()
{
    m_Type = Type::Number;
    m_Number = static_cast<double>(value);
    m_Integral = true;
    m_Magnitude = value;
}

JsonValue::JsonValue(double value) : JsonValue()
{
    m_Type = Type::Number;
    m_Number = value;
}

JsonValue::JsonValue(const char* value) : JsonValue(std::string(value)) {}

JsonValue::JsonValue(const std::string& value) : JsonValue()
{
    m_Type = Type::String;
    m_String = value;
}

JsonValue JsonValue::MakeArray()
{
    JsonValue value;
    value.m_Type = Type::Array;
    return value;
}

JsonValue JsonValue::MakeObject()
{
    JsonValue value;
    value.m_Type = Type::Object;
    return value;
}

bool JsonValue::AsBool() const
{
    if (m_Type != Type::Bool)
    {
        throw KttException("JSON value is not a boolean");
    }
    return m_Bool;
}

double JsonValue::AsDouble() const
{
    if (m_Type != Type::Number)
    {
        throw KttException("JSON value is not a number");
    }
    return m_Number;
}

int64_t JsonValue::AsInt() const
{
    if (m_Type != Type::Number)
    {
        throw KttException("JSON value is not a number");
    }
    if (m_Integral)
    {
        return m_Negative ? -static_cast<int64_t>(m_Magnitude - 1) - 1 : static_cast<int64_t>(m_Magnitude);
    }
    return static_cast<int64_t>(m_Number);
}

uint64_t JsonValue::AsUint() const
{
    if (m_Type != Type::Number)
    {
        throw KttException("JSON value is not a number");
    }
    if (m_Integral && !m_Negative)
    {
        return m_Magnitude;
    }
    return static_cast<uint64_t>(m_Number);
}

const std::string& JsonValue::AsString() const
{
    if (m_Type != Type::String)
    {
        throw KttException("JSON value is not a string");
    }
    return m_String;
}

void JsonValue::PushBack(const JsonValue& value)
{
    if (m_Type != Type::Array)
    {
        throw KttException("JSON value is not an array");
    }
    m_Array.push_back(value);
}

size_t JsonValue::Size() const
{
    if (m_Type == Type::Array)
    {
        return m_Array.size();
    }
    if (m_Type == Type::Object)
    {
        return m_Object.size();
    }
    return 0;
}

const JsonValue& JsonValue::operator[](const size_t index) const
{
    if (m_Type != Type::Array || index >= m_Array.size())
    {
        throw KttException("Invalid JSON array access");
    }
    return m_Array[index];
}

JsonValue& JsonValue::operator[](const std::string& key)
{
    if (m_Type == Type::Null)
    {
        m_Type = Type::Object;
    }
    if (m_Type != Type::Object)
    {
        throw KttException("JSON value is not an object");
    }
    for (auto& [name, value] : m_Object)
    {
        if (name == key)
        {
            return value;
        }
    }
    m_Object.emplace_back(key, JsonValue());
    return m_Object.back().second;
}

const JsonValue& JsonValue::At(const std::string& key) const
{
    if (m_Type != Type::Object)
    {
        throw KttException("JSON value is not an object");
    }
    for (const auto& [name, value] : m_Object)
    {
        if (name == key)
        {
            return value;
        }
    }
    throw KttException("JSON object has no key " + key);
}

bool JsonValue::Contains(const std::string& key) const
{
    if (m_Type != Type::Object)
    {
        return false;
    }
    for (const auto& [name, value] : m_Object)
    {
        if (name == key)
        {
            return true;
        }
    }
    return false;
}

static void DumpString(std::string& out, const std::string& value)
{
    out += '"';
    for (const char c : value)
    {
        switch (c)
        {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default:
            if (static_cast<unsigned char>(c) < 0x20)
            {
                char buffer[8];
                std::snprintf(buffer, sizeof(buffer), "\\u%04x", static_cast<unsigned char>(c));
                out += buffer;
            }
            else
            {
                out += c;
            }
        }
    }
    out += '"';
}

void JsonValue::DumpTo(std::string& out, const int indent, const int level) const
{
    const std::string pad = indent > 0 ? std::string(static_cast<size_t>(indent * (level + 1)), ' ') : "";
    const std::string closePad = indent > 0 ? std::string(static_cast<size_t>(indent * level), ' ') : "";
    const char* newline = indent > 0 ? "\n" : "";
    switch (m_Type)
    {
    case Type::Null:
        out += "null";
        break;
    case Type::Bool:
        out += m_Bool ? "true" : "false";
        break;
    case Type::Number:
        if (m_Integral)
        {
            out += m_Negative ? std::to_string(AsInt()) : std::to_string(m_Magnitude);
        }
        else if (!std::isfinite(m_Number))
        {
            out += "null";
        }
        else
        {
            char buffer[40];
            std::snprintf(buffer, sizeof(buffer), "%.17g", m_Number);
            std::string text = buffer;
            if (text.find_first_of(".eE") == std::string::npos)
            {
                text += ".0";
            }
            out += text;
        }
        break;
    case Type::String:
        DumpString(out, m_String);
        break;
    case Type::Array:
        if (m_Array.empty())
        {
            out += "[]";
            break;
        }
        out += "[";
        out += newline;
        for (size_t i = 0; i < m_Array.size(); ++i)
        {
            out += pad;
            m_Array[i].DumpTo(out, indent, level + 1);
            if (i + 1 < m_Array.size())
            {
                out += ",";
            }
            out += newline;
        }
        out += closePad + "]";
        break;
    case Type::Object:
        if (m_Object.empty())
        {
            out += "{}";
            break;
        }
        out += "{";
        out += newline;
        for (size_t i = 0; i < m_Object.size(); ++i)
        {
            out += pad;
            DumpString(out, m_Object[i].first);
            out += indent > 0 ? ": " : ":";
            m_Object[i].second.DumpTo(out, indent, level + 1);
            if (i + 1 < m_Object.size())
            {
                out += ",";
            }
            out += newline;
        }
        out += closePad + "}";
        break;
    }
}

std::string JsonValue::Dump(const int indent) const
{
    std::string out;
    DumpTo(out, indent, 0);
    return out;
}

namespace
{

class Parser
{
public:
    explicit Parser(const std::string& text) : m_Text(text), m_Position(0) {}

    JsonValue ParseDocument()
    {
        JsonValue value = ParseValue();
        SkipWhitespace();
        if (m_Position != m_Text.size())
        {
            Fail("trailing characters");
        }
        return value;
    }

private:
    const std::string& m_Text;
    size_t m_Position;

    [[noreturn]] void Fail(const std::string& message) const
    {
        throw KttException("JSON parse error at offset " + std::to_string(m_Position) + ": " + message);
    }

    void SkipWhitespace()
    {
        while (m_Position < m_Text.size() && (m_Text[m_Position] == ' ' || m_Text[m_Position] == '\n' || m_Text[m_Position] == '\r'
            || m_Text[m_Position] == '\t'))
        {
            ++m_Position;
        }
    }

    char Peek()
    {
        SkipWhitespace();
        if (m_Position >= m_Text.size())
        {
            Fail("unexpected end of input");
        }
        return m_Text[m_Position];
    }

    void Expect(const char c)
    {
        if (Peek() != c)
        {
            Fail(std::string("expected '") + c + "'");
        }
        ++m_Position;
    }

    bool Consume(const std::string& word)
    {
        if (m_Text.compare(m_Position, word.size(), word) == 0)
        {
            m_Position += word.size();
            return true;
        }
        return false;
    }

    JsonValue ParseValue()
    {
        const char c = Peek();
        if (c == '{')
        {
            return ParseObject();
        }
        if (c == '[')
        {
            return ParseArray();
        }
        if (c == '"')
        {
            return JsonValue(ParseString());
        }
        if (Consume("true"))
        {
            return JsonValue(true);
        }
        if (Consume("false"))
        {
            return JsonValue(false);
        }
        if (Consume("null"))
        {
            return JsonValue();
        }
        return ParseNumber();
    }

    JsonValue ParseObject()
    {
        Expect('{');
        JsonValue object = JsonValue::MakeObject();
        if (Peek() == '}')
        {
            ++m_Position;
            return object;
        }
        while (true)
        {
            if (Peek() != '"')
            {
                Fail("expected object key");
            }
            const std::string key = ParseString();
            Expect(':');
            object[key] = ParseValue();
            const char next = Peek();
            ++m_Position;
            if (next == '}')
            {
                return object;
            }
            if (next != ',')
            {
                Fail("expected ',' or '}'");
            }
        }
    }

    JsonValue ParseArray()
    {
        Expect('[');
        JsonValue array = JsonValue::MakeArray();
        if (Peek() == ']')
        {
            ++m_Position;
            return array;
        }
        while (true)
        {
            array.PushBack(ParseValue());
            const char next = Peek();
            ++m_Position;
            if (next == ']')
            {
                return array;
            }
            if (next != ',')
            {
                Fail("expected ',' or ']'");
            }
        }
    }

    static void AppendUtf8(std::string& out, const uint32_t code)
    {
        if (code < 0x80)
        {
            out += static_cast<char>(code);
        }
        else if (code < 0x800)
        {
            out += static_cast<char>(0xC0 | (code >> 6));
            out += static_cast<char>(0x80 | (code & 0x3F));
        }
        else
        {
            out += static_cast<char>(0xE0 | (code >> 12));
            out += static_cast<char>(0x80 | ((code >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (code & 0x3F));
        }
    }

    std::string ParseString()
    {
        Expect('"');
        std::string result;
        while (true)
        {
            if (m_Position >= m_Text.size())
            {
                Fail("unterminated string");
            }
            const char c = m_Text[m_Position++];
            if (c == '"')
            {
                return result;
            }
            if (c != '\\')
            {
                result += c;
                continue;
            }
            if (m_Position >= m_Text.size())
            {
                Fail("unterminated escape");
            }
            const char e = m_Text[m_Position++];
            switch (e)
            {
            case '"': result += '"'; break;
            case '\\': result += '\\'; break;
            case '/': result += '/'; break;
            case 'b': result += '\b'; break;
            case 'f': result += '\f'; break;
            case 'n': result += '\n'; break;
            case 'r': result += '\r'; break;
            case 't': result += '\t'; break;
            case 'u':
            {
                if (m_Position + 4 > m_Text.size())
                {
                    Fail("invalid unicode escape");
                }
                const uint32_t code = static_cast<uint32_t>(std::strtoul(m_Text.substr(m_Position, 4).c_str(), nullptr, 16));
                m_Position += 4;
                AppendUtf8(result, code);
                break;
            }
            default:
                Fail("invalid escape");
            }
        }
    }

    JsonValue ParseNumber()
    {
        SkipWhitespace();
        const size_t start = m_Position;
        if (m_Position < m_Text.size() && (m_Text[m_Position] == '-' || m_Text[m_Position] == '+'))
        {
            ++m_Position;
        }
        bool integral = true;
        while (m_Position < m_Text.size())
        {
            const char c = m_Text[m_Position];
            if (c >= '0' && c <= '9')
            {
                ++m_Position;
            }
            else if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+')
            {
                integral = false;
                ++m_Position;
            }
            else
            {
                break;
            }
        }
        const std::string token = m_Text.substr(start, m_Position - start);
        if (token.empty() || token == "-" || token == "+")
        {
            Fail("invalid value");
        }
        if (integral)
        {
            errno = 0;
            if (token[0] == '-')
            {
                const long long value = std::strtoll(token.c_str(), nullptr, 10);
                if (errno == 0)
                {
                    return JsonValue(static_cast<int64_t>(value));
                }
            }
            else
            {
                const unsigned long long value = std::strtoull(token.c_str(), nullptr, 10);
                if (errno == 0)
                {
                    return JsonValue(static_cast<uint64_t>(value));
                }
            }
        }
        return JsonValue(std::strtod(token.c_str(), nullptr));
    }
};

} // namespace

JsonValue JsonValue::Parse(const std::string& text)
{
    Parser parser(text);
    return parser.ParseDocument();
}

} // namespace ktt
