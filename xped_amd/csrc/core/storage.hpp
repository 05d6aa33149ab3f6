// Storage wire format of a tensor (storage.cpp): the fields of the reference's StorageType
// serialisation in one binary layout.
#pragma once
#include <string>
#include <vector>

#include "tensor.hpp"

namespace xa {

std::vector<char> serialize(const DTensor& t);
DTensor deserialize(Runtime* rt, const char* buf, std::size_t n);
void save(const DTensor& t, const std::string& path);
DTensor load(Runtime* rt, const std::string& path);

}  // namespace xa
