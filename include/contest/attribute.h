// Contest column data types.
//
// Mirrors the contest API's `DataType` (reference include/attribute.h:8-13). The
// enumerator order is part of the ABI: plans and tables built against the
// contest header pass these values through unchanged.
#pragma once

#include <cstdint>
#include <string>

enum class DataType {
    INT32,   // 4-byte integer
    INT64,   // 8-byte integer
    FP64,    // 8-byte IEEE double
    VARCHAR, // string of arbitrary length
};

struct Attribute {
    DataType    type;
    std::string name;
};
