/** @file DimensionVector.h
  * Three-dimensional size of a kernel launch (global or local).
  * Same operations as Source/Api/Configuration/DimensionVector.cpp:60-263 of the
  * reference (component-wise Multiply, Divide, RoundUp, DivideCeil, modification
  * by one value in one dimension).
  */
#pragma once

#include <cstddef>
#include <string>
#include <vector>

namespace ktt
{

enum class ModifierAction;
enum class ModifierDimension;

class DimensionVector
{
public:
    DimensionVector();
    explicit DimensionVector(const size_t sizeX);
    explicit DimensionVector(const size_t sizeX, const size_t sizeY);
    explicit DimensionVector(const size_t sizeX, const size_t sizeY, const size_t sizeZ);
    explicit DimensionVector(const std::vector<size_t>& vector);

    void SetSizeX(const size_t sizeX);
    void SetSizeY(const size_t sizeY);
    void SetSizeZ(const size_t sizeZ);
    void SetSize(const ModifierDimension dimension, const size_t size);

    void Multiply(const DimensionVector& factor);
    void Divide(const DimensionVector& divisor);
    void RoundUp(const DimensionVector& multiple);
    void DivideCeil(const DimensionVector& divisor);
    void ModifyByValue(const size_t value, const ModifierAction action, const ModifierDimension dimension);

    size_t GetSizeX() const;
    size_t GetSizeY() const;
    size_t GetSizeZ() const;
    size_t GetSize(const ModifierDimension dimension) const;
    size_t GetTotalSize() const;
    std::vector<size_t> GetVector() const;
    std::string GetString() const;

    bool operator==(const DimensionVector& other) const;
    bool operator!=(const DimensionVector& other) const;

private:
    size_t m_SizeX;
    size_t m_SizeY;
    size_t m_SizeZ;
};

} // namespace ktt
