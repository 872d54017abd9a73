// molden.hpp -- Molden orbital output (scf::writeOrbitalsMolden,
// src/chem/scf/WriteOrbitals.cpp:17-137) and the number text it prints.
#pragma once
#include <string>

#include "scf.hpp"

namespace flo {

// The fields of flo::NumberFormat (src/util/FormattedStream.hpp:19-34) that
// decide a number's text.
struct NumberFormat {
    unsigned max_characters = 16;
    int significant_digits = 8;
    bool adjust_positives = true;  // a space where a minus sign would go
    bool truncate_zeros = false;
    bool scientific_notation = true;
    unsigned max_mantissa = 4;
    static NumberFormat scientific(unsigned width, int digits);                    // scientificFormat
    static NumberFormat crude(unsigned width, int digits, bool adjust_positives = true);  // crudeFormat(Positive)
};

std::string formatNumber(double x, const NumberFormat& f);     // toFormattedString(double), FormattedStream.cpp:91-183
std::string formatInteger(long long x, const NumberFormat& f);  // toFormattedString(i64) for non-scientific formats

}  // namespace flo

namespace scf {

// [Atoms] in bohr, [GTO] with the basis file's contraction coefficients,
// [MO] with every orbital (or the occupied ones) in Molden's AO order.
int writeOrbitalsMolden(const Job& job, const std::string& path, bool write_virtual, std::string* error);

}  // namespace scf
