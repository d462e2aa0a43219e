/** @file KttException.h
  * Exception thrown inside the library; public Tuner methods catch it, log it and
  * return Invalid* values (Source/Tuner.cpp:41-52 pattern). The reason maps to a
  * ResultStatus in the kernel runner (Source/KernelRunner/KernelRunner.cpp:329-343).
  */
#pragma once

#include <stdexcept>
#include <string>

#include <KttTypes.h>

namespace ktt
{

class KttException : public std::runtime_error
{
public:
    explicit KttException(const std::string& message, const ExceptionReason reason = ExceptionReason::General) :
        std::runtime_error(message),
        m_Reason(reason)
    {}

    ExceptionReason GetReason() const
    {
        return m_Reason;
    }

private:
    ExceptionReason m_Reason;
};

} // namespace ktt
