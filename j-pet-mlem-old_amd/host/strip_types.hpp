// Host-side types with the shapes the reference's strip code uses, so that
// include/strip_gpu_reconstruction.hpp is called the same way from this
// project's CLI as from the reference's reconstruction_cmd.cpp.
//   PET2D::Pixel<S>            src/2d/geometry/pixel.h:11-59 (subset)
//   PET2D::PixelMap<P, V>      src/2d/geometry/pixel_map.h:15-162 (subset)
//   PET2D::Strip::Response<F>  src/2d/strip/response.h:13-45 (data layout)
//   PET2D::Strip::Scanner<F,S> src/2d/strip/scanner.h:26-291 (constructor
//                              arguments and the members the GPU entry reads)
#pragma once

#include <cstring>
#include <memory>
#include <type_traits>

namespace PET2D {

template <typename SType>
struct Pixel {
  using S = SType;
  using Size = typename std::common_type<S, int>::type;
  Pixel() = default;
  Pixel(S x, S y) : x(x), y(y) {}
  S x = 0, y = 0;
  Size index(S width) const { return static_cast<Size>(y) * width + x; }
};

/// Row-major image, index y * width + x.  Owning or a view over caller data.
template <typename PixelType, typename ValueType>
class PixelMap {
 public:
  using Pixel = PixelType;
  using S = typename Pixel::S;
  using Size = typename Pixel::Size;
  using Value = ValueType;

  PixelMap(S width, S height)
      : size(static_cast<Size>(width) * height),
        width(width),
        height(height),
        storage_(new Value[size]()),
        data(storage_.get()) {}
  PixelMap(S width, S height, const Value& value) : PixelMap(width, height) {
    assign(value);
  }
  /// Non-owning view (pixel_map.h:35-41)
  PixelMap(S width, S height, Value* data)
      : size(static_cast<Size>(width) * height),
        width(width),
        height(height),
        data(data) {}
  PixelMap(const PixelMap&) = delete;
  PixelMap& operator=(const PixelMap&) = delete;

  Value& operator[](Size i) { return data[i]; }
  const Value& operator[](Size i) const { return data[i]; }
  void assign(const Value& v) {
    for (Size i = 0; i < size; ++i) data[i] = v;
  }
  Value* begin() { return data; }
  Value* end() { return data + size; }
  const Value* begin() const { return data; }
  const Value* end() const { return data + size; }

  const Size size;
  const S width;
  const S height;

 private:
  std::unique_ptr<Value[]> storage_;

 public:
  Value* const data;
};

namespace Strip {

template <typename FType>
struct Response {
  using F = FType;
  F z_u, z_d, dl;
  Response() = default;
  Response(F z_u, F z_d, F dl) : z_u(z_u), z_d(z_d), dl(dl) {}
};

template <typename FType, typename SType>
class Scanner {
 public:
  using F = FType;
  using S = SType;
  /// Same argument order as scanner.h:39-49
  Scanner(F radius, F scintillator_length, S n_y_pixels, S n_z_pixels,
          F pixel_height, F pixel_width, F sigma_z, F sigma_dl,
          F center_y = 0, F center_z = 0)
      : radius(radius),
        scintillator_length(scintillator_length),
        n_y_pixels(n_y_pixels),
        n_z_pixels(n_z_pixels),
        total_n_pixels(static_cast<int>(n_y_pixels) * n_z_pixels),
        pixel_width(pixel_width),
        pixel_height(pixel_height),
        sigma_z(sigma_z),
        sigma_dl(sigma_dl),
        center_y(center_y),
        center_z(center_z) {}

  const F radius;
  const F scintillator_length;
  const S n_y_pixels;
  const S n_z_pixels;
  const int total_n_pixels;
  const F pixel_width;
  const F pixel_height;
  const F sigma_z;
  const F sigma_dl;
  const F center_y;
  const F center_z;
};

}  // namespace Strip
}  // namespace PET2D
