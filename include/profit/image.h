/*
 * profit/image.h -- 2-D coordinates and the Image / Mask containers of the
 * C++ host API with libprofit's layout and semantics
 * (include/profit/image.h:47-823, src/image.cpp): row-major storage,
 * element (x, y) at index x + y * width, unsigned dimensions.
 */
#ifndef PROFIT_AMD_IMAGE_H
#define PROFIT_AMD_IMAGE_H

#include <algorithm>
#include <ostream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace profit {

/* A point or a size on a 2-D grid (libprofit's _2dcoordinate). */
class _2dcoordinate {
public:
	_2dcoordinate() : x(0), y(0) {}
	_2dcoordinate(unsigned int x, unsigned int y) : x(x), y(y) {}
	unsigned int x, y;

	bool operator==(const _2dcoordinate &o) const { return x == o.x && y == o.y; }
	bool operator!=(const _2dcoordinate &o) const { return !(*this == o); }
	bool operator<(const _2dcoordinate &o) const { return x < o.x && y < o.y; }
	bool operator<=(const _2dcoordinate &o) const { return x <= o.x && y <= o.y; }
	bool operator>(const _2dcoordinate &o) const { return x > o.x && y > o.y; }
	bool operator>=(const _2dcoordinate &o) const { return x >= o.x && y >= o.y; }
	explicit operator bool() const { return x != 0 || y != 0; }
	_2dcoordinate operator+(const _2dcoordinate &o) const { return {x + o.x, y + o.y}; }
	_2dcoordinate operator-(const _2dcoordinate &o) const { return {x - o.x, y - o.y}; }
	_2dcoordinate operator*(unsigned int f) const { return {x * f, y * f}; }
	_2dcoordinate operator/(unsigned int f) const { return {x / f, y / f}; }
	_2dcoordinate &operator+=(const _2dcoordinate &o) { x += o.x; y += o.y; return *this; }
	_2dcoordinate &operator-=(const _2dcoordinate &o) { x -= o.x; y -= o.y; return *this; }
	_2dcoordinate &operator*=(unsigned int f) { x *= f; y *= f; return *this; }
	_2dcoordinate &operator/=(unsigned int f) { x /= f; y /= f; return *this; }
};

inline std::ostream &operator<<(std::ostream &os, const _2dcoordinate &c)
{
	return os << "[" << c.x << ", " << c.y << "]";
}

typedef _2dcoordinate Point;
typedef _2dcoordinate Dimensions;
typedef std::pair<Point, Point> PointPair;
typedef std::pair<double, double> PixelScale;

/* A rectangle [first, second) (libprofit's Box). */
class Box : public PointPair {
public:
	Box() : PointPair() {}
	Box(Point lo, Point hi) : PointPair(lo, hi)
	{
		if (!(lo <= hi)) throw std::invalid_argument("box's lower boundary must be <= than high boundary");
	}
	bool empty() const { return first == second; }
};

/* Common part of Image and Mask. */
template <typename T, typename D>
class surface {
public:
	typedef typename std::vector<T>::reference reference;
	typedef typename std::vector<T>::const_reference const_reference;
	typedef typename std::vector<T>::iterator iterator;
	typedef typename std::vector<T>::const_iterator const_iterator;

	surface() = default;
	explicit surface(Dimensions d) : dims(d), data_(std::size_t(d.x) * d.y) {}
	surface(T value, Dimensions d) : dims(d), data_(std::size_t(d.x) * d.y, value) {}
	surface(const std::vector<T> &v, Dimensions d) : dims(d), data_(v) { check(); }
	surface(std::vector<T> &&v, Dimensions d) : dims(d), data_(std::move(v)) { check(); }

	unsigned int getWidth() const { return dims.x; }
	unsigned int getHeight() const { return dims.y; }
	Dimensions getDimensions() const { return dims; }
	unsigned int size() const { return (unsigned int)data_.size(); }
	bool empty() const { return data_.empty(); }
	explicit operator bool() const { return !empty(); }
	reference operator[](std::size_t i) { return data_[i]; }
	const_reference operator[](std::size_t i) const { return data_[i]; }
	reference operator[](const Point &p) { return data_[p.x + std::size_t(p.y) * dims.x]; }
	const_reference operator[](const Point &p) const { return data_[p.x + std::size_t(p.y) * dims.x]; }
	bool operator==(const surface &o) const { return dims == o.dims && data_ == o.data_; }
	bool operator!=(const surface &o) const { return !(*this == o); }
	iterator begin() { return data_.begin(); }
	iterator end() { return data_.end(); }
	const_iterator begin() const { return data_.begin(); }
	const_iterator end() const { return data_.end(); }
	void zero() { std::fill(data_.begin(), data_.end(), T(0)); }
	std::vector<T> &_get() { return data_; }
	const std::vector<T> &_get() const { return data_; }
	explicit operator std::vector<T>() const { return data_; }

	/* copy of this surface placed at `start` of a zero surface of `new_dims`
	 * (include/profit/image.h:399-425) */
	D extend(Dimensions new_dims, Point start = Point()) const
	{
		if (new_dims.x < dims.x) throw std::invalid_argument("new_width should be >= width");
		if (new_dims.y < dims.y) throw std::invalid_argument("new_height should be >= height");
		if (start.x + dims.x > new_dims.x) throw std::invalid_argument("start_x + new_width should be <= width");
		if (start.y + dims.y > new_dims.y) throw std::invalid_argument("start_y + new_height <= image.height");
		D out(new_dims);
		for (unsigned int j = 0; j < dims.y; j++)
			for (unsigned int i = 0; i < dims.x; i++) out[Point(i + start.x, j + start.y)] = (*this)[Point(i, j)];
		return out;
	}
	/* the `new_dims` block starting at `start` (include/profit/image.h:433-448) */
	D crop(Dimensions new_dims, Point start) const
	{
		if (new_dims.x > dims.x) throw std::invalid_argument("new_width should be <= width");
		if (new_dims.y > dims.y) throw std::invalid_argument("new_height should be <= height");
		if (start.x + new_dims.x > dims.x) throw std::invalid_argument("start_x + new_width should be <= image.width");
		if (start.y + new_dims.y > dims.y) throw std::invalid_argument("start_y + new_height should be <= image.height");
		D out(new_dims);
		for (unsigned int j = 0; j < new_dims.y; j++)
			for (unsigned int i = 0; i < new_dims.x; i++) out[Point(i, j)] = (*this)[Point(i + start.x, j + start.y)];
		return out;
	}
	/* element order reversed (a 180 degree rotation) */
	D reverse() const
	{
		D out(static_cast<const D &>(*this));
		std::reverse(out.data_.begin(), out.data_.end());
		return out;
	}
	/* smallest box holding every non-zero element (include/profit/image.h:468-501) */
	Box bounding_box() const
	{
		Point lo = dims, hi;
		bool any = false;
		for (unsigned int j = 0; j < dims.y; j++)
			for (unsigned int i = 0; i < dims.x; i++) {
				if ((*this)[Point(i, j)] == T(0)) continue;
				any = true;
				lo.x = std::min(lo.x, i);
				lo.y = std::min(lo.y, j);
				hi.x = std::max(hi.x, i + 1);
				hi.y = std::max(hi.y, j + 1);
			}
		if (!any) return Box();
		return Box(lo, hi);
	}

protected:
	void check() const
	{
		if (std::size_t(dims.x) * dims.y != data_.size()) throw std::invalid_argument("data.size() != weight * height");
	}
	Dimensions dims;
	std::vector<T> data_;
};

class Mask : public surface<bool, Mask> {
public:
	Mask() = default;
	Mask(unsigned int w, unsigned int h) : surface(Dimensions(w, h)) {}
	explicit Mask(Dimensions d) : surface(d) {}
	Mask(bool v, unsigned int w, unsigned int h) : surface(v, Dimensions(w, h)) {}
	Mask(bool v, Dimensions d) : surface(v, d) {}
	Mask(const std::vector<bool> &v, unsigned int w, unsigned int h) : surface(v, Dimensions(w, h)) {}
	Mask(const std::vector<bool> &v, Dimensions d) : surface(v, d) {}
	Mask(std::vector<bool> &&v, unsigned int w, unsigned int h) : surface(std::move(v), Dimensions(w, h)) {}
	Mask(std::vector<bool> &&v, Dimensions d) : surface(std::move(v), d) {}

	/* every element within `pad` (a (2 pad.x + 1) x (2 pad.y + 1) box) of a
	 * set element is set (src/image.cpp:112-162) */
	Mask expand_by(Dimensions pad, int threads = 1) const;
	/* each element becomes a factor x factor block (src/image.cpp:164-182) */
	Mask upsample(unsigned int factor) const;
	/* one byte per element (the C-ABI layout) */
	std::vector<unsigned char> bytes() const;
};

class Image : public surface<double, Image> {
public:
	enum DownsamplingMode { SUM = 0, AVERAGE, SAMPLE };
	enum UpsamplingMode { SCALE = 0, COPY };  // SCALE divides the values by factor^2

	Image() = default;
	Image(unsigned int w, unsigned int h) : surface(Dimensions(w, h)) {}
	explicit Image(Dimensions d) : surface(d) {}
	Image(double v, unsigned int w, unsigned int h) : surface(v, Dimensions(w, h)) {}
	Image(double v, Dimensions d) : surface(v, d) {}
	Image(const std::vector<double> &v, unsigned int w, unsigned int h) : surface(v, Dimensions(w, h)) {}
	Image(const std::vector<double> &v, Dimensions d) : surface(v, d) {}
	Image(std::vector<double> &&v, unsigned int w, unsigned int h) : surface(std::move(v), Dimensions(w, h)) {}
	Image(std::vector<double> &&v, Dimensions d) : surface(std::move(v), d) {}

	double total() const;
	/* divided by its sum if the sum is > 0 (src/image.cpp:208-221) */
	Image normalize() const;
	void normalize();
	Image upsample(unsigned int factor, UpsamplingMode mode = SCALE) const;
	Image downsample(unsigned int factor, DownsamplingMode mode = SUM) const;

	Image &operator+=(const Image &o);
	Image &operator-=(const Image &o);
	Image &operator*=(const Image &o);
	Image &operator/=(const Image &o);
	Image &operator+=(double v);
	Image &operator-=(double v);
	Image &operator*=(double v);
	Image &operator/=(double v);
	Image operator+(const Image &o) const { Image r(*this); r += o; return r; }
	Image operator-(const Image &o) const { Image r(*this); r -= o; return r; }
	Image operator*(const Image &o) const { Image r(*this); r *= o; return r; }
	Image operator/(const Image &o) const { Image r(*this); r /= o; return r; }
	Image operator+(double v) const { Image r(*this); r += v; return r; }
	Image operator-(double v) const { Image r(*this); r -= v; return r; }
	Image operator*(double v) const { Image r(*this); r *= v; return r; }
	Image operator/(double v) const { Image r(*this); r /= v; return r; }
	/* zero where the mask is not set */
	Image &operator&=(const Mask &m);
	Image operator&(const Mask &m) const { Image r(*this); r &= m; return r; }
};

}  // namespace profit

#endif
