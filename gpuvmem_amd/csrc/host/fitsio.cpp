// FITS image writer (include/gvm/fitsio.h): one primary HDU, 80-character
// header cards in 2880-byte records, BITPIX -32 big-endian data padded with
// zeros to a whole record (the FITS 4.0 standard, sections 3.3, 4.1, 5.3).
#include "gvm/fitsio.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gvm_hip.h"

namespace {
thread_local std::string g_fits_error;

int fail(const std::string& msg) {
  g_fits_error = msg;
  return 1;
}

// One 80-character card: keyword (8, left-justified), "= ", value, comment.
std::string card(const std::string& key, const std::string& value, const std::string& comment) {
  std::string c = key;
  c.resize(8, ' ');
  if (!value.empty() || key == "END") {
    if (key != "END") c += "= " + value;
  }
  if (!comment.empty()) c += " / " + comment;
  c.resize(80, ' ');
  return c;
}

std::string fixed_str(const char* s) {
  // quoted string value, padded to at least 8 characters inside the quotes
  std::string v = s ? s : "";
  std::string q;
  for (char ch : v) q += (ch == '\'') ? std::string("''") : std::string(1, ch);
  while (q.size() < 8) q += ' ';
  return "'" + q + "'";
}

std::string fixed_num(double x) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%20.12E", x);
  return buf;
}

std::string fixed_int(long x) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%20ld", x);
  return buf;
}

std::string fixed_logical(bool b) {
  std::string v(19, ' ');
  return v + (b ? "T" : "F");
}

}  // namespace

extern "C" const char* gvm_fits_last_error(void) { return g_fits_error.c_str(); }

extern "C" int gvm_fits_write_image(const char* path, const float* image, int on_device, long M,
                                    long N, int index, float fg_scale,
                                    const gvm_fits_header* hdr) {
  if (!path || !image || !hdr || M <= 0 || N <= 0 || index < 0)
    return fail("gvm_fits_write_image: bad argument");
  const long n = M * N;
  std::vector<float> data(n);
  if (on_device) {
    if (gvm_memcpy_d2h(data.data(), image + n * index, sizeof(float) * n) != GVM_OK)
      return fail("gvm_fits_write_image: device copy failed");
  } else {
    std::memcpy(data.data(), image + n * index, sizeof(float) * n);
  }
  // OCopyFITS scales by fg_scale on output (internal units -> Jy/pixel)
  for (long k = 0; k < n; k++) data[k] *= fg_scale;

  std::vector<std::string> cards;
  const bool freq_axis = hdr->cdelt3 != 0.0;
  cards.push_back(card("SIMPLE", fixed_logical(true), "conforms to FITS standard"));
  cards.push_back(card("BITPIX", fixed_int(-32), "IEEE single precision"));
  cards.push_back(card("NAXIS", fixed_int(freq_axis ? 3 : 2), ""));
  cards.push_back(card("NAXIS1", fixed_int(M), ""));
  cards.push_back(card("NAXIS2", fixed_int(N), ""));
  if (freq_axis) cards.push_back(card("NAXIS3", fixed_int(1), ""));
  cards.push_back(card("BSCALE", fixed_num(1.0), ""));
  cards.push_back(card("BZERO", fixed_num(0.0), ""));
  cards.push_back(card("BUNIT", fixed_str(hdr->bunit), "Unit of measurement"));
  cards.push_back(card("NITER", fixed_int(hdr->niter), "Number of iteration in gpuvmem software"));
  if (hdr->object) cards.push_back(card("OBJECT", fixed_str(hdr->object), ""));
  cards.push_back(card("RADESYS", fixed_str(hdr->radesys ? hdr->radesys : "ICRS"), ""));
  cards.push_back(card("EQUINOX", fixed_num(hdr->equinox), ""));
  cards.push_back(card("CTYPE1", fixed_str("RA---SIN"), ""));
  cards.push_back(card("CRPIX1", fixed_num(hdr->crpix1), ""));
  cards.push_back(card("CRVAL1", fixed_num(hdr->crval1), ""));
  cards.push_back(card("CDELT1", fixed_num(hdr->cdelt1), ""));
  cards.push_back(card("CUNIT1", fixed_str("deg"), ""));
  cards.push_back(card("CTYPE2", fixed_str("DEC--SIN"), ""));
  cards.push_back(card("CRPIX2", fixed_num(hdr->crpix2), ""));
  cards.push_back(card("CRVAL2", fixed_num(hdr->crval2), ""));
  cards.push_back(card("CDELT2", fixed_num(hdr->cdelt2), ""));
  cards.push_back(card("CUNIT2", fixed_str("deg"), ""));
  if (freq_axis) {
    cards.push_back(card("CTYPE3", fixed_str("FREQ"), ""));
    cards.push_back(card("CRPIX3", fixed_num(1.0), ""));
    cards.push_back(card("CRVAL3", fixed_num(hdr->crval3), ""));
    cards.push_back(card("CDELT3", fixed_num(hdr->cdelt3), ""));
    cards.push_back(card("CUNIT3", fixed_str("Hz"), ""));
  }
  cards.push_back(card("END", "", ""));

  std::string header;
  for (const auto& c : cards) header += c;
  header.resize(((header.size() + 2879) / 2880) * 2880, ' ');

  std::vector<unsigned char> body(((size_t)n * 4 + 2879) / 2880 * 2880, 0);
  for (long k = 0; k < n; k++) {
    uint32_t u;
    std::memcpy(&u, &data[k], 4);
    body[4 * k] = (unsigned char)(u >> 24);
    body[4 * k + 1] = (unsigned char)(u >> 16);
    body[4 * k + 2] = (unsigned char)(u >> 8);
    body[4 * k + 3] = (unsigned char)u;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(std::string("gvm_fits_write_image: cannot open ") + path);
  const bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size() &&
                  std::fwrite(body.data(), 1, body.size(), f) == body.size();
  if (std::fclose(f) != 0 || !ok) return fail(std::string("gvm_fits_write_image: write failed: ") + path);
  return 0;
}
