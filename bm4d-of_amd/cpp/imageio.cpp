// Image files of the command-line tools: PNG (zlib), uncompressed TIFF and
// Middlebury .flo, as planar float images (data[c*w*h + y*w + x]), the
// layout iio_read_image_float_split / iio_save_image_float_split give the
// reference (lib/iio/iio.c:3242, 3870).  Written for this project; only the
// file formats are shared.
//
// Conventions kept from the reference's tools:
//   * writing PNG quantises each sample as (uint8)clamp(v + 0.5, 0, 255)
//     (iio.c:737 with its T8 crop, iio.c:668);
//   * .flo files hold 2 channels, interleaved (u, v) per pixel after the tag
//     202021.25 and the int32 width and height (iio.c:1885-1900, 2742-2757).
#include "../../include/bm4dof_io.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string g_err;

// Sample type of the file, which decides how a gray conversion rounds.
enum class SampleKind { U8, U16, F32 };

bool fail(const std::string &m)
{
	g_err = m;
	return false;
}

bool read_file(const char *path, std::vector<unsigned char> &buf)
{
	FILE *f = std::fopen(path, "rb");
	if (!f) return fail(std::string(path) + ": cannot open");
	std::fseek(f, 0, SEEK_END);
	const long n = std::ftell(f);
	std::fseek(f, 0, SEEK_SET);
	buf.resize(n > 0 ? (size_t)n : 0);
	const size_t got = n > 0 ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
	std::fclose(f);
	if ((long)got != n) return fail(std::string(path) + ": short read");
	return true;
}

bool write_file(const char *path, const std::vector<unsigned char> &buf)
{
	FILE *f = std::fopen(path, "wb");
	if (!f) return fail(std::string(path) + ": cannot create");
	const size_t put = std::fwrite(buf.data(), 1, buf.size(), f);
	const bool ok = std::fclose(f) == 0 && put == buf.size();
	return ok ? true : fail(std::string(path) + ": write failed");
}

bool has_suffix(const std::string &s, const char *suf)
{
	const size_t n = std::strlen(suf);
	if (s.size() < n) return false;
	for (size_t i = 0; i < n; i++)
		if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
	return true;
}

// ---------------------------------------------------------------- PNG

uint32_t be32(const unsigned char *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
void put_be32(std::vector<unsigned char> &b, uint32_t v)
{
	b.push_back((unsigned char)(v >> 24));
	b.push_back((unsigned char)(v >> 16));
	b.push_back((unsigned char)(v >> 8));
	b.push_back((unsigned char)v);
}

int paeth(int a, int b, int c)
{
	const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
	return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

bool read_png(const std::vector<unsigned char> &f, std::vector<float> &out, int &w, int &h, int &c,
              SampleKind &kind)
{
	if (f.size() < 33) return fail("png: truncated");
	size_t o = 8;
	int depth = 0, ctype = -1, interlace = 0;
	std::vector<unsigned char> idat;
	while (o + 12 <= f.size()) {
		const uint32_t len = be32(&f[o]);
		const std::string type(reinterpret_cast<const char *>(&f[o + 4]), 4);
		if (o + 12 + len > f.size()) return fail("png: truncated chunk");
		const unsigned char *d = &f[o + 8];
		if (type == "IHDR") {
			if (len != 13) return fail("png: bad IHDR chunk");
			w = (int)be32(d);
			h = (int)be32(d + 4);
			depth = d[8];
			ctype = d[9];
			interlace = d[12];
		} else if (type == "IDAT")
			idat.insert(idat.end(), d, d + len);
		else if (type == "IEND")
			break;
		o += 12 + len;
	}
	if (w <= 0 || h <= 0) return fail("png: bad size");
	if (interlace) return fail("png: interlaced files are not supported");
	if (depth != 8 && depth != 16) return fail("png: only 8- and 16-bit samples are supported");
	switch (ctype) {
	case 0: c = 1; break;
	case 2: c = 3; break;
	case 4: c = 2; break;
	case 6: c = 4; break;
	default: return fail("png: palette images are not supported");
	}
	const int bps = depth / 8, bpp = bps * c;
	// the sample limit of the reference's image library: fewer than 2^32
	if ((uint64_t)w * (uint64_t)h * (uint64_t)c >= (1ull << 32)) return fail("png: image too large");
	kind = bps == 1 ? SampleKind::U8 : SampleKind::U16;
	const size_t stride = (size_t)w * bpp;
	std::vector<unsigned char> raw((stride + 1) * h);
	uLongf rawlen = (uLongf)raw.size();
	if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size())
		return fail("png: corrupt image data");
	std::vector<unsigned char> prev(stride, 0), cur(stride);
	out.assign((size_t)w * h * c, 0.f);
	for (int y = 0; y < h; y++) {
		const unsigned char ft = raw[(stride + 1) * y];
		const unsigned char *s = &raw[(stride + 1) * y + 1];
		for (size_t i = 0; i < stride; i++) {
			const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev[i], cc = i >= (size_t)bpp ? prev[i - bpp] : 0;
			int v = s[i];
			switch (ft) {
			case 0: break;
			case 1: v += a; break;
			case 2: v += b; break;
			case 3: v += (a + b) / 2; break;
			case 4: v += paeth(a, b, cc); break;
			default: return fail("png: bad filter");
			}
			cur[i] = (unsigned char)v;
		}
		for (int x = 0; x < w; x++)
			for (int k = 0; k < c; k++) {
				const unsigned char *px = &cur[(size_t)x * bpp + (size_t)k * bps];
				const float v = bps == 1 ? (float)px[0] : (float)((px[0] << 8) | px[1]);
				out[(size_t)k * w * h + (size_t)y * w + x] = v;
			}
		std::swap(prev, cur);
	}
	return true;
}

void png_chunk(std::vector<unsigned char> &b, const char *type, const unsigned char *d, size_t n)
{
	put_be32(b, (uint32_t)n);
	const size_t start = b.size();
	b.insert(b.end(), type, type + 4);
	if (n) b.insert(b.end(), d, d + n);
	put_be32(b, (uint32_t)crc32(0, &b[start], (uInt)(4 + n)));
}

bool write_png(const char *path, const float *data, int w, int h, int c)
{
	if (c < 1 || c > 4) return fail("png: 1 to 4 channels");
	const int ctype = c == 1 ? 0 : c == 2 ? 4 : c == 3 ? 2 : 6;
	const size_t stride = (size_t)w * c;
	std::vector<unsigned char> raw((stride + 1) * h);
	for (int y = 0; y < h; y++) {
		raw[(stride + 1) * y] = 0;
		for (int x = 0; x < w; x++)
			for (int k = 0; k < c; k++) {
				const float v = 0.5f + data[(size_t)k * w * h + (size_t)y * w + x];
				const float t = v > 0.f ? (v < 255.f ? v : 255.f) : 0.f;
				raw[(stride + 1) * y + 1 + (size_t)x * c + k] = (unsigned char)t;
			}
	}
	uLongf zl = compressBound((uLong)raw.size());
	std::vector<unsigned char> z(zl);
	if (compress2(z.data(), &zl, raw.data(), (uLong)raw.size(), 6) != Z_OK) return fail("png: deflate failed");
	std::vector<unsigned char> b = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
	unsigned char ihdr[13];
	const uint32_t ww = (uint32_t)w, hh = (uint32_t)h;
	for (int i = 0; i < 4; i++) {
		ihdr[i] = (unsigned char)(ww >> (24 - 8 * i));
		ihdr[4 + i] = (unsigned char)(hh >> (24 - 8 * i));
	}
	ihdr[8] = 8;
	ihdr[9] = (unsigned char)ctype;
	ihdr[10] = ihdr[11] = ihdr[12] = 0;
	png_chunk(b, "IHDR", ihdr, 13);
	png_chunk(b, "IDAT", z.data(), zl);
	png_chunk(b, "IEND", nullptr, 0);
	return write_file(path, b);
}

// ---------------------------------------------------------------- TIFF

struct TiffReader {
	const std::vector<unsigned char> &f;
	bool le;
	uint32_t u16(size_t o) const { return le ? (f[o] | f[o + 1] << 8) : (f[o] << 8 | f[o + 1]); }
	uint32_t u32(size_t o) const
	{
		return le ? (uint32_t)(f[o] | f[o + 1] << 8 | f[o + 2] << 16 | (uint32_t)f[o + 3] << 24)
		          : (uint32_t)((uint32_t)f[o] << 24 | f[o + 1] << 16 | f[o + 2] << 8 | f[o + 3]);
	}
};

bool read_tiff(const std::vector<unsigned char> &f, std::vector<float> &out, int &w, int &h, int &c,
               SampleKind &kind)
{
	if (f.size() < 8) return fail("tiff: truncated");
	TiffReader r{f, f[0] == 'I'};
	size_t ifd = r.u32(4);
	if (ifd + 2 > f.size()) return fail("tiff: bad directory");
	const uint32_t n = r.u16(ifd);
	uint32_t bits = 8, comp = 1, spp = 1, rps = 0xffffffffu, planar = 1, fmt = 1;
	std::vector<uint32_t> offs, cnts;
	w = h = 0;
	auto values = [&](size_t e, std::vector<uint32_t> &v) {
		const uint32_t type = r.u16(e + 2), count = r.u32(e + 4);
		const size_t sz = type == 3 ? 2 : 4;
		const size_t base = count * sz <= 4 ? e + 8 : r.u32(e + 8);
		v.resize(count);
		for (uint32_t i = 0; i < count; i++) v[i] = type == 3 ? r.u16(base + i * 2) : r.u32(base + i * 4);
	};
	for (uint32_t i = 0; i < n; i++) {
		const size_t e = ifd + 2 + 12 * (size_t)i;
		const uint32_t tag = r.u16(e);
		std::vector<uint32_t> v;
		values(e, v);
		if (v.empty()) continue;
		switch (tag) {
		case 256: w = (int)v[0]; break;
		case 257: h = (int)v[0]; break;
		case 258: bits = v[0]; break;
		case 259: comp = v[0]; break;
		case 273: offs = v; break;
		case 277: spp = v[0]; break;
		case 278: rps = v[0]; break;
		case 279: cnts = v; break;
		case 284: planar = v[0]; break;
		case 339: fmt = v[0]; break;
		default: break;
		}
	}
	if (w <= 0 || h <= 0 || offs.empty()) return fail("tiff: missing size or strips");
	if (comp != 1) return fail("tiff: only uncompressed files are supported");
	if (!((bits == 8 && fmt == 1) || (bits == 16 && fmt == 1) || (bits == 32 && fmt == 3)))
		return fail("tiff: samples must be 8- or 16-bit unsigned or 32-bit float");
	c = (int)spp;
	const size_t bps = bits / 8;
	kind = bits == 8 ? SampleKind::U8 : bits == 16 ? SampleKind::U16 : SampleKind::F32;
	// concatenate the strips
	std::vector<unsigned char> pix;
	for (size_t s = 0; s < offs.size(); s++) {
		const size_t len = s < cnts.size() ? cnts[s] : 0;
		if (offs[s] + len > f.size()) return fail("tiff: strip outside the file");
		pix.insert(pix.end(), f.begin() + offs[s], f.begin() + offs[s] + len);
	}
	(void)rps;
	const size_t nsamp = (size_t)w * h * c;
	if (pix.size() < nsamp * bps) return fail("tiff: not enough pixel data");
	out.assign(nsamp, 0.f);
	for (size_t i = 0; i < nsamp; i++) {
		const size_t o = i * bps;
		float v;
		if (bps == 1) v = pix[o];
		else if (bps == 2) v = (float)(r.le ? (pix[o] | pix[o + 1] << 8) : (pix[o] << 8 | pix[o + 1]));
		else {
			uint32_t u = r.le ? (uint32_t)(pix[o] | pix[o + 1] << 8 | pix[o + 2] << 16 | (uint32_t)pix[o + 3] << 24)
			                  : (uint32_t)((uint32_t)pix[o] << 24 | pix[o + 1] << 16 | pix[o + 2] << 8 | pix[o + 3]);
			std::memcpy(&v, &u, 4);
		}
		// chunky: sample i is (pixel i / c, channel i % c); planar: channel-major
		size_t dst;
		if (planar == 2) dst = i;
		else dst = (i % c) * (size_t)w * h + i / c;
		out[dst] = v;
	}
	return true;
}

void le16(std::vector<unsigned char> &b, uint32_t v)
{
	b.push_back((unsigned char)v);
	b.push_back((unsigned char)(v >> 8));
}
void le32(std::vector<unsigned char> &b, uint32_t v)
{
	for (int i = 0; i < 4; i++) b.push_back((unsigned char)(v >> (8 * i)));
}

bool write_tiff(const char *path, const float *data, int w, int h, int c)
{
	// float32, chunky, one strip, little endian
	const size_t nsamp = (size_t)w * h * c;
	std::vector<unsigned char> b = {'I', 'I', 42, 0};
	const uint32_t data_off = 8;
	const uint32_t data_len = (uint32_t)(nsamp * 4);
	const uint32_t bits_off = data_off + data_len;               // BitsPerSample array (c > 2)
	const uint32_t fmt_off = bits_off + (c > 2 ? 2 * c : 0);     // SampleFormat array (c > 2)
	uint32_t ifd_off = fmt_off + (c > 2 ? 2 * c : 0);
	ifd_off += ifd_off & 1;
	le32(b, ifd_off);
	for (int y = 0; y < h; y++)
		for (int x = 0; x < w; x++)
			for (int k = 0; k < c; k++) {
				uint32_t u;
				std::memcpy(&u, &data[(size_t)k * w * h + (size_t)y * w + x], 4);
				le32(b, u);
			}
	if (c > 2) {
		for (int k = 0; k < c; k++) le16(b, 32);
		for (int k = 0; k < c; k++) le16(b, 3);
	}
	while (b.size() < ifd_off) b.push_back(0);
	struct E { uint16_t tag, type; uint32_t count, value; };
	std::vector<E> es;
	es.push_back({256, 4, 1, (uint32_t)w});
	es.push_back({257, 4, 1, (uint32_t)h});
	es.push_back({258, 3, (uint32_t)c, c > 2 ? bits_off : 32u});
	es.push_back({259, 3, 1, 1});
	es.push_back({262, 3, 1, c >= 3 ? 2u : 1u});
	es.push_back({273, 4, 1, data_off});
	es.push_back({277, 3, 1, (uint32_t)c});
	es.push_back({278, 4, 1, (uint32_t)h});
	es.push_back({279, 4, 1, data_len});
	es.push_back({284, 3, 1, 1});
	es.push_back({339, 3, (uint32_t)c, c > 2 ? fmt_off : 3u});
	le16(b, (uint32_t)es.size());
	for (const E &e : es) {
		le16(b, e.tag);
		le16(b, e.type);
		le32(b, e.count);
		if (e.type == 3 && e.count <= 2) {
			// inline shorts: first value in the low half; two shorts both packed
			const uint32_t v = (e.count == 2 && (e.tag == 258 || e.tag == 339)) ? (e.value | e.value << 16) : e.value;
			le32(b, v);
		} else
			le32(b, e.value);
	}
	le32(b, 0);
	return write_file(path, b);
}

// ---------------------------------------------------------------- FLO

bool read_flo(const std::vector<unsigned char> &f, std::vector<float> &out, int &w, int &h, int &c)
{
	if (f.size() < 12) return fail("flo: truncated");
	float tag;
	int32_t ww, hh;
	std::memcpy(&tag, &f[0], 4);
	std::memcpy(&ww, &f[4], 4);
	std::memcpy(&hh, &f[8], 4);
	if (tag != 202021.25f || ww <= 0 || hh <= 0) return fail("flo: bad header");
	w = ww;
	h = hh;
	c = 2;
	const size_t n = (size_t)w * h;
	if (f.size() < 12 + n * 8) return fail("flo: truncated data");
	out.resize(2 * n);
	for (size_t i = 0; i < n; i++) {
		std::memcpy(&out[i], &f[12 + 8 * i], 4);
		std::memcpy(&out[n + i], &f[16 + 8 * i], 4);
	}
	return true;
}

bool write_flo(const char *path, const float *data, int w, int h, int c)
{
	if (c != 2) return fail("flo: needs exactly 2 channels");
	const size_t n = (size_t)w * h;
	std::vector<unsigned char> b(12 + 8 * n);
	const float tag = 202021.25f;
	const int32_t ww = w, hh = h;
	std::memcpy(&b[0], &tag, 4);
	std::memcpy(&b[4], &ww, 4);
	std::memcpy(&b[8], &hh, 4);
	for (size_t i = 0; i < n; i++) {
		std::memcpy(&b[12 + 8 * i], &data[i], 4);
		std::memcpy(&b[16 + 8 * i], &data[n + i], 4);
	}
	return write_file(path, b);
}

bool read_any(const char *path, std::vector<float> &img, int &w, int &h, int &c, SampleKind &kind)
{
	std::vector<unsigned char> f;
	if (!read_file(path, f)) return false;
	if (f.size() >= 8 && f[0] == 0x89 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G')
		return read_png(f, img, w, h, c, kind);
	if (f.size() >= 4 && ((f[0] == 'I' && f[1] == 'I' && f[2] == 42) || (f[0] == 'M' && f[1] == 'M' && f[3] == 42)))
		return read_tiff(f, img, w, h, c, kind);
	if (f.size() >= 4 && f[0] == 'P' && f[1] == 'I' && f[2] == 'E' && f[3] == 'H') {
		kind = SampleKind::F32;
		return read_flo(f, img, w, h, c);
	}
	return fail(std::string(path) + ": unknown format (PNG, TIFF and .flo are read)");
}

int to_malloc(const std::vector<float> &img, float **data)
{
	*data = (float *)std::malloc(img.size() * sizeof(float));
	if (!*data) return fail("out of memory"), -1;
	std::memcpy(*data, img.data(), img.size() * sizeof(float));
	return 0;
}

} // namespace

extern "C" {

__attribute__((visibility("default")))
const char *bm4dof_io_error(void) { return g_err.c_str(); }

__attribute__((visibility("default")))
int bm4dof_io_read(const char *path, float **data, int *w, int *h, int *c)
{
	if (!path || !data || !w || !h || !c) return fail("NULL argument"), -1;
	std::vector<float> img;
	SampleKind kind;
	if (!read_any(path, img, *w, *h, *c, kind)) return -1;
	return to_malloc(img, data);
}

// Gray as iio_read_image_float gives it (lib/iio/iio.c:3486-3505 with
// iio_hacky_uncolorize, :935-973): 3 or 4 channels become
// .299 R + .587 G + .114 B computed in double and stored back in the file's
// sample type (integer samples truncate); 2 channels are an error.
__attribute__((visibility("default")))
int bm4dof_io_read_gray(const char *path, float **data, int *w, int *h)
{
	if (!path || !data || !w || !h) return fail("NULL argument"), -1;
	std::vector<float> img;
	SampleKind kind;
	int c;
	if (!read_any(path, img, *w, *h, c, kind)) return -1;
	if (c == 2 || c > 4) return fail(std::string(path) + ": non-scalarizable image"), -1;
	if (c == 4 && kind == SampleKind::U16) return fail(std::string(path) + ": 16-bit RGBA is not converted to gray"), -1;
	if (c >= 3) {
		const size_t n = (size_t)*w * *h;
		std::vector<float> g(n);
		for (size_t i = 0; i < n; i++) {
			const double y = .299 * img[i] + .587 * img[n + i] + .114 * img[2 * n + i];
			g[i] = kind == SampleKind::U8    ? (float)(uint8_t)y
			       : kind == SampleKind::U16 ? (float)(uint16_t)y
			                                 : (float)y;
		}
		img.swap(g);
	}
	return to_malloc(img, data);
}

__attribute__((visibility("default")))
int bm4dof_io_write(const char *path, const float *data, int w, int h, int c)
{
	if (!path || !data || w <= 0 || h <= 0 || c <= 0) return fail("bad arguments"), -1;
	const std::string p(path);
	bool ok;
	if (has_suffix(p, ".png")) ok = write_png(path, data, w, h, c);
	else if (has_suffix(p, ".flo")) ok = write_flo(path, data, w, h, c);
	else if (has_suffix(p, ".tif") || has_suffix(p, ".tiff")) ok = write_tiff(path, data, w, h, c);
	else ok = fail(p + ": unknown extension (.png, .tif, .tiff and .flo are written)");
	return ok ? 0 : -1;
}

__attribute__((visibility("default")))
void bm4dof_io_free(float *data) { std::free(data); }

} // extern "C"
