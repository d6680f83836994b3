// format.cpp -- the reference's fmt "{}" rendering of doubles
// (output/sinks.cpp:100-204, pinned by tests/cpp/lennardjonesium/output/test_sinks.cpp:100-225):
// shortest round-trip digits; fixed notation when the leading digit's decimal
// exponent lies in [-4, 16), scientific "d.ddde+XX" otherwise; integral values
// print without a decimal point.

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "lj_host.h"

namespace lj_host {

std::string format_double(double v)
{
    if (std::isnan(v)) return "nan";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    char sci[64];
    auto res = std::to_chars(sci, sci + sizeof(sci), v, std::chars_format::scientific);
    std::string s(sci, res.ptr);
    // split "[-]d[.ddd]e[+-]XX"
    bool neg = false;
    size_t p = 0;
    if (s[0] == '-') { neg = true; p = 1; }
    const size_t epos = s.find('e');
    std::string digits;
    for (size_t k = p; k < epos; ++k)
        if (s[k] != '.') digits += s[k];
    const int exp = std::stoi(s.substr(epos + 1));
    std::string out = neg ? "-" : "";
    if (exp < -4 || exp >= 16) {
        out += digits[0];
        if (digits.size() > 1) out += "." + digits.substr(1);
        char eb[16];
        std::snprintf(eb, sizeof(eb), "e%c%02d", exp < 0 ? '-' : '+', exp < 0 ? -exp : exp);
        return out + eb;
    }
    if (exp < 0) {
        out += "0.";
        out += std::string(-exp - 1, '0');
        out += digits;
        return out;
    }
    const int intlen = exp + 1;
    if ((int)digits.size() <= intlen) {
        out += digits + std::string(intlen - digits.size(), '0');
        return out;
    }
    out += digits.substr(0, intlen) + "." + digits.substr(intlen);
    return out;
}

}  // namespace lj_host

extern "C" int lj_format_double(double value, char* buf, int cap)
{
    const std::string s = lj_host::format_double(value);
    if (buf && cap > 0) {
        const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return (int)s.size();
}
