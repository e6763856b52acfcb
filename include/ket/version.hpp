// Version of the ket API this backend re-declares (reference VERSION).
#pragma once

#include <string_view>

namespace ket {

inline constexpr std::string_view version = "0.1.0";

}  // namespace ket
