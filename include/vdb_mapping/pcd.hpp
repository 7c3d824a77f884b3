// PCD files of the drop-in persistence calls (saveMapToPCD / loadMapFromPCD):
// a writer of PCL's ASCII layout for x y z float points (header as
// pcl::PCDWriter::generateHeader writes it for pcl::PointXYZ, values with 8
// significant digits, one point per line) and a reader of ASCII and binary
// (uncompressed) PCD files with float x, y and z fields. Host code only.
#ifndef VDB_MAPPING_PCD_HPP
#define VDB_MAPPING_PCD_HPP

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace vdb_mapping {
namespace pcd {

// xyz: 3 floats per point. Returns false when the file cannot be written.
inline bool writeAscii(const std::string& path, const std::vector<float>& xyz)
{
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f)
  {
    return false;
  }
  const size_t n = xyz.size() / 3;
  std::fprintf(f,
               "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z\nSIZE 4 4 4\n"
               "TYPE F F F\nCOUNT 1 1 1\nWIDTH %zu\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS %zu\n"
               "DATA ascii\n",
               n,
               n);
  for (size_t i = 0; i < n; ++i)
  {
    for (int a = 0; a < 3; ++a)
    {
      const float v = xyz[3 * i + a];
      if (std::isnan(v))
      {
        std::fputs("nan", f);
      }
      else
      {
        std::fprintf(f, "%.8g", (double)v);
      }
      std::fputc(a < 2 ? ' ' : '\n', f);
    }
  }
  return std::fclose(f) == 0;
}

// Reads the x, y, z float fields of every point into xyz. Returns false on a
// missing file, a malformed header, missing or non-float x/y/z fields, or a
// compressed data section.
inline bool read(const std::string& path, std::vector<float>& xyz)
{
  std::ifstream in(path, std::ios::binary);
  if (!in)
  {
    return false;
  }
  std::vector<std::string> fields, types;
  std::vector<int> sizes, counts;
  long points = -1, width = -1, height = 1;
  std::string data, line;
  while (std::getline(in, line))
  {
    if (line.empty() || line[0] == '#')
    {
      continue;
    }
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    if (key == "FIELDS")
    {
      for (std::string v; ls >> v;) fields.push_back(v);
    }
    else if (key == "SIZE")
    {
      for (int v; ls >> v;) sizes.push_back(v);
    }
    else if (key == "TYPE")
    {
      for (std::string v; ls >> v;) types.push_back(v);
    }
    else if (key == "COUNT")
    {
      for (int v; ls >> v;) counts.push_back(v);
    }
    else if (key == "WIDTH")
    {
      ls >> width;
    }
    else if (key == "HEIGHT")
    {
      ls >> height;
    }
    else if (key == "POINTS")
    {
      ls >> points;
    }
    else if (key == "DATA")
    {
      ls >> data;
      break;
    }
  }
  if (counts.empty())
  {
    counts.assign(fields.size(), 1);
  }
  if (points < 0)
  {
    points = width * height;
  }
  if (fields.empty() || sizes.size() != fields.size() || types.size() != fields.size() ||
      counts.size() != fields.size() || points < 0)
  {
    return false;
  }
  int col[3] = {-1, -1, -1};
  size_t off[3] = {0, 0, 0}, stride = 0;
  int values = 0;  // values per point (ASCII columns)
  for (size_t k = 0; k < fields.size(); ++k)
  {
    for (int a = 0; a < 3; ++a)
    {
      if (fields[k] == std::string(1, "xyz"[a]))
      {
        if (types[k] != "F" || sizes[k] != 4)
        {
          return false;
        }
        col[a] = values;
        off[a] = stride;
      }
    }
    values += counts[k];
    stride += (size_t)sizes[k] * counts[k];
  }
  if (col[0] < 0 || col[1] < 0 || col[2] < 0)
  {
    return false;
  }
  xyz.assign((size_t)points * 3, 0.0f);
  if (data == "ascii")
  {
    for (long i = 0; i < points; ++i)
    {
      if (!std::getline(in, line))
      {
        return false;
      }
      std::istringstream ls(line);
      std::vector<std::string> tok;
      for (std::string v; ls >> v;) tok.push_back(v);
      if ((int)tok.size() < values)
      {
        return false;
      }
      for (int a = 0; a < 3; ++a)
      {
        xyz[3 * i + a] = std::strtof(tok[col[a]].c_str(), nullptr);
      }
    }
    return true;
  }
  if (data == "binary")
  {
    std::vector<char> buf(stride);
    for (long i = 0; i < points; ++i)
    {
      if (!in.read(buf.data(), (std::streamsize)stride))
      {
        return false;
      }
      for (int a = 0; a < 3; ++a)
      {
        std::memcpy(&xyz[3 * i + a], buf.data() + off[a], 4);
      }
    }
    return true;
  }
  return false;  // binary_compressed is not read
}

} // namespace pcd
} // namespace vdb_mapping

#endif // VDB_MAPPING_PCD_HPP
