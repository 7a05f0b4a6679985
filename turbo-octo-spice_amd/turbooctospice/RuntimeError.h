// RuntimeError.h -- the library's exception type (mirrors the reference's
// turbooctospice/RuntimeError.h:8-18).
#ifndef TOS_AMD_TURBOOCTOSPICE_RUNTIME_ERROR
#define TOS_AMD_TURBOOCTOSPICE_RUNTIME_ERROR

#include <stdexcept>
#include <string>

namespace turbooctospice {
class RuntimeError : public std::runtime_error {
public:
    explicit RuntimeError(std::string const &reason) : std::runtime_error(reason) {}
};
}  // namespace turbooctospice

#endif
