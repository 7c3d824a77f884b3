// Types the drop-in API shares with the reference: pcl::PointXYZ /
// pcl::PointCloud, Eigen::Matrix<double,3,1>, openvdb::Coord / Vec3d.
//
// The real PCL, Eigen and OpenVDB headers are used when they are found
// (__has_include). Otherwise minimal stand-ins with the same names, member
// names and memory layout are declared, enough for the occupancy-integration
// API (points[i].x/y/z at a 16-byte stride, origin.x()/y()/z(), Coord(x,y,z),
// accessor getValue / isValueOn). Define VDBGPU_NO_EXTERNAL_TYPES to force the
// stand-ins.
#ifndef VDB_MAPPING_COMPAT_HPP
#define VDB_MAPPING_COMPAT_HPP

#include <climits>
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

#if !defined(VDBGPU_NO_EXTERNAL_TYPES) && __has_include(<pcl/point_types.h>) && \
    __has_include(<pcl/point_cloud.h>)
#include <pcl/point_cloud.h>
#include <pcl/point_types.h>
#define VDBGPU_HAVE_PCL 1
#else
namespace pcl {
struct PointXYZ {
  float x, y, z;
  float data_pad = 1.0f;  // pcl::PointXYZ is 16 bytes (x, y, z, padding)
  PointXYZ() : x(0), y(0), z(0) {}
  PointXYZ(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};
template <typename PointT>
struct PointCloud {
  using Ptr = std::shared_ptr<PointCloud<PointT>>;
  using ConstPtr = std::shared_ptr<const PointCloud<PointT>>;
  std::vector<PointT> points;
  uint32_t width = 0, height = 1;
  size_t size() const { return points.size(); }
  bool empty() const { return points.empty(); }
  void push_back(const PointT& p) { points.push_back(p); }
  typename std::vector<PointT>::const_iterator begin() const { return points.begin(); }
  typename std::vector<PointT>::const_iterator end() const { return points.end(); }
};
}  // namespace pcl
static_assert(sizeof(pcl::PointXYZ) == 16, "PointXYZ stand-in must keep the 16-byte stride");
#endif

#if !defined(VDBGPU_NO_EXTERNAL_TYPES) && __has_include(<eigen3/Eigen/Core>)
#include <eigen3/Eigen/Core>
#define VDBGPU_HAVE_EIGEN 1
#else
namespace Eigen {
template <typename Scalar, int Rows, int Cols>
class Matrix;
template <>
class Matrix<double, 3, 1> {
 public:
  Matrix() : v_{0, 0, 0} {}
  Matrix(double x, double y, double z) : v_{x, y, z} {}
  double x() const { return v_[0]; }
  double y() const { return v_[1]; }
  double z() const { return v_[2]; }
  double& operator()(int i) { return v_[i]; }
  double operator()(int i) const { return v_[i]; }
  const double* data() const { return v_; }

 private:
  double v_[3];
};
template <>
class Matrix<double, 4, 1> {
 public:
  Matrix() : v_{0, 0, 0, 0} {}
  Matrix(double x, double y, double z, double w) : v_{x, y, z, w} {}
  double x() const { return v_[0]; }
  double y() const { return v_[1]; }
  double z() const { return v_[2]; }
  double w() const { return v_[3]; }
  double& operator()(int i) { return v_[i]; }
  double operator()(int i) const { return v_[i]; }

 private:
  double v_[4];
};
template <>
class Matrix<double, 4, 4> {
 public:
  Matrix() : m_{} {}
  static Matrix Identity() {
    Matrix r;
    for (int i = 0; i < 4; ++i) r.m_[i][i] = 1.0;
    return r;
  }
  double& operator()(int r, int c) { return m_[r][c]; }
  double operator()(int r, int c) const { return m_[r][c]; }

 private:
  double m_[4][4];
};
}  // namespace Eigen
#endif

#if !defined(VDBGPU_NO_EXTERNAL_TYPES) && __has_include(<openvdb/openvdb.h>)
#include <openvdb/openvdb.h>
#define VDBGPU_HAVE_OPENVDB 1
#else
namespace openvdb {
using Index64 = uint64_t;
class Coord {
 public:
  Coord() : v_{0, 0, 0} {}
  Coord(int32_t x, int32_t y, int32_t z) : v_{x, y, z} {}
  int32_t x() const { return v_[0]; }
  int32_t y() const { return v_[1]; }
  int32_t z() const { return v_[2]; }
  int32_t operator[](int i) const { return v_[i]; }
  bool operator==(const Coord& o) const { return v_[0] == o.v_[0] && v_[1] == o.v_[1] && v_[2] == o.v_[2]; }
  bool operator!=(const Coord& o) const { return !(*this == o); }

 private:
  int32_t v_[3];
};
namespace math {
template <typename T>
class Vec3 {
 public:
  Vec3() : v_{0, 0, 0} {}
  Vec3(T x, T y, T z) : v_{x, y, z} {}
  T x() const { return v_[0]; }
  T y() const { return v_[1]; }
  T z() const { return v_[2]; }
  T& x() { return v_[0]; }
  T& y() { return v_[1]; }
  T& z() { return v_[2]; }
  T operator[](int i) const { return v_[i]; }
  T& operator[](int i) { return v_[i]; }

 private:
  T v_[3];
};
}  // namespace math
using Vec3d = math::Vec3<double>;
class CoordBBox {
 public:
  CoordBBox() : min_(INT32_MAX, INT32_MAX, INT32_MAX), max_(INT32_MIN, INT32_MIN, INT32_MIN) {}
  CoordBBox(const Coord& mn, const Coord& mx) : min_(mn), max_(mx) {}
  const Coord& min() const { return min_; }
  const Coord& max() const { return max_; }
  bool isInside(const Coord& c) const {
    for (int a = 0; a < 3; ++a)
      if (c[a] < min_[a] || c[a] > max_[a]) return false;
    return true;
  }

 private:
  Coord min_, max_;
};
class BBoxd {
 public:
  BBoxd() = default;
  BBoxd(const Vec3d& mn, const Vec3d& mx) : min_(mn), max_(mx) {}
  const Vec3d& min() const { return min_; }
  const Vec3d& max() const { return max_; }

 private:
  Vec3d min_, max_;
};
}  // namespace openvdb
#endif

#endif  // VDB_MAPPING_COMPAT_HPP
