// Mesh files in the reference's formats (gmsh 2 ASCII and its binary form).
#pragma once
#include <string>

#include "element_stream.hpp"

namespace s355 {

ElementStream read_gmsh_ascii(const std::string& path);
ElementStream read_gmsh_binary(const std::string& path);
void write_gmsh_ascii(const ElementStream& es, const std::string& path);
void write_gmsh_binary(const ElementStream& es, const std::string& path);

}  // namespace s355
