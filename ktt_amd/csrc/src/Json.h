// Minimal JSON document model (parse / dump) for tuning results and tuned-configuration files.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace ktt
{

class JsonValue
{
public:
    enum class Type
    {
        Null,
        Bool,
        Number,
        String,
        Array,
        Object
    };

    JsonValue();
    JsonValue(std::nullptr_t);
    JsonValue(bool value);
    JsonValue(int value);
    JsonValue(int64_t value);
    JsonValue(uint64_t value);
    JsonValue(double value);
    JsonValue(const char* value);
    JsonValue(const std::string& value);

    static JsonValue MakeArray();
    static JsonValue MakeObject();
    static JsonValue Parse(const std::string& text);

    Type GetType() const { return m_Type; }
    bool IsNull() const { return m_Type == Type::Null; }
    bool IsNumber() const { return m_Type == Type::Number; }
    bool IsString() const { return m_Type == Type::String; }
    bool IsArray() const { return m_Type == Type::Array; }
    bool IsObject() const { return m_Type == Type::Object; }
    bool IsBool() const { return m_Type == Type::Bool; }
    /** Number written without fraction or exponent, and within integer range. */
    bool IsInteger() const { return m_Type == Type::Number && m_Integral; }
    bool IsNegative() const { return m_Type == Type::Number && m_Negative; }

    bool AsBool() const;
    double AsDouble() const;
    int64_t AsInt() const;
    uint64_t AsUint() const;
    const std::string& AsString() const;

    // Array access
    void PushBack(const JsonValue& value);
    size_t Size() const;
    const JsonValue& operator[](const size_t index) const;

    // Object access (keys kept in insertion order for stable output)
    JsonValue& operator[](const std::string& key);
    const JsonValue& At(const std::string& key) const;
    bool Contains(const std::string& key) const;
    const std::vector<std::pair<std::string, JsonValue>>& Items() const { return m_Object; }

    std::string Dump(const int indent = 2) const;

private:
    Type m_Type;
    bool m_Bool;
    double m_Number;
    bool m_Integral;
    bool m_Negative;
    uint64_t m_Magnitude;
    std::string m_String;
    std::vector<JsonValue> m_Array;
    std::vector<std::pair<std::string, JsonValue>> m_Object;

    void DumpTo(std::string& out, const int indent, const int level) const;
};

} // namespace ktt
