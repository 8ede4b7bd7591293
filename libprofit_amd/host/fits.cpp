// FITS image I/O (include/profit/fits.h).  The FITS standard fixes the
// layout: 80-character header cards in 2880-byte blocks ending with END,
// then the data as big-endian IEEE doubles, padded to a whole block.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "profit/fits.h"

namespace profit {

namespace {

constexpr std::size_t BLOCK = 2880;

double big_endian(double v)
{
	uint64_t u;
	std::memcpy(&u, &v, 8);
	u = __builtin_bswap64(u);
	std::memcpy(&v, &u, 8);
	return v;
}

void card(std::ostream &f, const std::string &text)
{
	std::string c = text.substr(0, 80);
	c.resize(80, ' ');
	f.write(c.data(), 80);
}

std::string keyword(const char *name, const std::string &value)
{
	char buf[96];
	std::snprintf(buf, sizeof(buf), "%-8s= %20s", name, value.c_str());
	return buf;
}

}  // namespace

Image from_fits(const std::string &filename, PixelScale &pixel_scale)
{
	std::ifstream f(filename, std::ios::binary);
	if (!f) throw invalid_file("Couldn't open '" + filename + "' for reading: " + std::strerror(errno));
	unsigned int w = 0, h = 0;
	int bitpix = 0, naxis = 0;
	double cx = 1, cy = 1;
	std::size_t read = 0;
	for (;;) {
		char c[81] = {0};
		f.read(c, 80);
		if (f.gcount() != 80) throw invalid_file("File " + filename + " does not look like a normal fits file");
		read += 80;
		std::string key(c, 8);
		key.erase(key.find_last_not_of(' ') + 1);
		const char *val = c[8] == '=' ? c + 9 : nullptr;
		if (key == "END") break;
		if (!val) continue;
		if (key == "NAXIS1") w = (unsigned int)std::strtoul(val, nullptr, 10);
		else if (key == "NAXIS2") h = (unsigned int)std::strtoul(val, nullptr, 10);
		else if (key == "NAXIS") naxis = (int)std::strtol(val, nullptr, 10);
		else if (key == "BITPIX") bitpix = (int)std::strtol(val, nullptr, 10);
		else if (key == "CDELT1") cx = std::strtod(val, nullptr);
		else if (key == "CDELT2") cy = std::strtod(val, nullptr);
	}
	if (bitpix != -64 || naxis != 2) {
		std::ostringstream os;
		os << "File " << filename << ": only 2-D BITPIX=-64 images are supported (BITPIX=" << bitpix
		   << ", NAXIS=" << naxis << ")";
		throw invalid_file(os.str());
	}
	f.seekg((std::streamoff)(((read + BLOCK - 1) / BLOCK) * BLOCK), std::ios::beg);
	std::vector<double> v((std::size_t)w * h);
	f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * 8));
	if ((std::size_t)f.gcount() != v.size() * 8) throw invalid_file("Error while reading file: less data found than expected");
	for (double &x : v) x = big_endian(x);
	pixel_scale = {cx, cy};
	return Image(std::move(v), w, h);
}

void to_fits(const Image &image, const Point &offset, const PixelScale &pixel_scale, std::string fname)
{
	if (fname.size() <= 5 || fname.compare(fname.size() - 5, 5, ".fits") != 0) fname += ".fits";
	std::ofstream f(fname, std::ios::binary);
	if (!f) throw invalid_file("Couldn't open '" + fname + "' for writing: " + std::strerror(errno));
	auto num = [](double v) {
		char b[32];
		std::snprintf(b, sizeof(b), "%.10G", v);
		return std::string(b);
	};
	card(f, keyword("SIMPLE", "T"));
	card(f, keyword("BITPIX", "-64"));
	card(f, keyword("NAXIS", "2"));
	card(f, keyword("NAXIS1", std::to_string(image.getWidth())));
	card(f, keyword("NAXIS2", std::to_string(image.getHeight())));
	card(f, keyword("CRPIX1", "1"));
	card(f, keyword("CRVAL1", num((0.5 - offset.x) * pixel_scale.first)));
	card(f, keyword("CDELT1", num(pixel_scale.first)));
	card(f, keyword("CRPIX2", "1"));
	card(f, keyword("CRVAL2", num((0.5 - offset.y) * pixel_scale.second)));
	card(f, keyword("CDELT2", num(pixel_scale.second)));
	card(f, "END");
	const std::size_t hdr = 12 * 80;
	f.write(std::string(((hdr + BLOCK - 1) / BLOCK) * BLOCK - hdr, ' ').data(),
	        (std::streamsize)(((hdr + BLOCK - 1) / BLOCK) * BLOCK - hdr));
	for (double v : image) {
		const double b = big_endian(v);
		f.write(reinterpret_cast<const char *>(&b), 8);
	}
	const std::size_t nb = (std::size_t)image.size() * 8;
	const std::size_t pad = ((nb + BLOCK - 1) / BLOCK) * BLOCK - nb;
	f.write(std::string(pad, '\0').data(), (std::streamsize)pad);
	if (!f) throw invalid_file("Error while writing " + fname);
}

}  // namespace profit
