// dada.cpp -- DADA ASCII header parsing for the raw-file source (host only).
//
// Restates the keyword search of ascii_header_find / ascii_header_get
// (Kernel/Classes/ascii_header.c) and the fields ASCIIObservation::load reads
// (Kernel/Classes/ASCIIObservation.C:24-45, 100-420), so that a DADA file's
// header configures the engines as the reference's DADAFile
// (Kernel/Classes/DADAFile.C:46-68) configures its Observation.
#include <cerrno>
#include <cstdlib>
#include <cstring>

#include "dspx_internal.hpp"

namespace dspx {

// The first occurrence of key that starts the header or a line (or follows
// a backslash) and is followed by a space or tab; the header's first word is
// taken as found whatever follows it.
static const char* header_find(const char* header, const char* key)
{
  const size_t n = std::strlen(key);
  const char* k = std::strstr(header, key);
  while (k && k > header) {
    const char before = k[-1], after = k[n];
    if ((before == '\n' || before == '\\') && (after == '\t' || after == ' ')) break;
    k = std::strstr(k + 1, key);
  }
  return k;
}

// The whitespace-delimited token after key (what sscanf "%s" reads).
static bool header_token(const char* header, const char* key, char* value, size_t cap)
{
  const char* k = header_find(header, key);
  if (!k || cap == 0) return false;
  const char* v = k + std::strcspn(k, " \t\n");
  v += std::strspn(v, " \t\n\r");
  const size_t len = std::strcspn(v, " \t\n\r");
  if (len == 0) return false;
  const size_t m = len < cap - 1 ? len : cap - 1;
  std::memcpy(value, v, m);
  value[m] = '\0';
  return true;
}

template <class T> static bool parse_number(const char* tok, T& out);

template <> bool parse_number<double>(const char* tok, double& out)
{
  char* end = nullptr;
  errno = 0;
  double v = std::strtod(tok, &end);
  if (end == tok || errno == ERANGE) return false;
  out = v;
  return true;
}

template <> bool parse_number<uint64_t>(const char* tok, uint64_t& out)
{
  char* end = nullptr;
  errno = 0;
  if (tok[0] == '-') return false;
  unsigned long long v = std::strtoull(tok, &end, 10);
  if (end == tok || errno == ERANGE) return false;
  out = v;
  return true;
}

template <> bool parse_number<int64_t>(const char* tok, int64_t& out)
{
  char* end = nullptr;
  errno = 0;
  long long v = std::strtoll(tok, &end, 10);
  if (end == tok || errno == ERANGE) return false;
  out = v;
  return true;
}

template <> bool parse_number<uint32_t>(const char* tok, uint32_t& out)
{
  uint64_t v;
  if (!parse_number<uint64_t>(tok, v) || v > 0xffffffffull) return false;
  out = uint32_t(v);
  return true;
}

}  // namespace dspx

using namespace dspx;

extern "C" {

int dspx_ascii_header_get(const char* header, const char* key, char* value, uint32_t cap)
{
  if (!header || !key || !value) return DSPX_E_PARAM;
  return header_token(header, key, value, cap) ? 1 : -1;
}

int dspx_dada_parse(const char* header, dspx_dada_info* out, char* missing)
{
  if (!header || !out) return DSPX_E_PARAM;
  *out = dspx_dada_info{};
  char tok[256];
  int rc = DSPX_OK;
  auto fail_key = [&](const char* key) {
    if (missing && rc == DSPX_OK) {
      std::strncpy(missing, key, 31);
      missing[31] = '\0';
    }
    rc = DSPX_E_PARAM;
  };
  // required keys: present and parsed, or the load fails (ASCIIObservation.h:133-142)
  auto req_str = [&](const char* key, char* dst, size_t cap) {
    if (!header_token(header, key, dst, cap)) fail_key(key);
  };
  auto req_num = [&](const char* key, auto& dst) {
    if (!header_token(header, key, tok, sizeof tok) || !parse_number(tok, dst)) fail_key(key);
  };
  auto opt_num = [&](const char* key, auto& dst, auto dflt) {
    if (!header_token(header, key, tok, sizeof tok) || !parse_number(tok, dst)) dst = dflt;
  };
  req_str("TELESCOPE", out->telescope, sizeof out->telescope);
  req_str("SOURCE", out->source, sizeof out->source);
  req_num("FREQ", out->centre_frequency);
  req_num("BW", out->bandwidth);
  opt_num("NCHAN", out->nchan, 1u);
  req_num("NPOL", out->npol);
  req_num("NBIT", out->nbit);
  opt_num("NDIM", out->ndim, 1u);
  opt_num("NDAT", out->ndat, uint64_t(0));
  req_num("TSAMP", out->tsamp);
  req_str("UTC_START", out->utc_start, sizeof out->utc_start);
  opt_num("PICOSECONDS", out->picoseconds, int64_t(0));
  // OBS_OFFSET, or OFFSET in older headers (ASCIIObservation.C:386-398)
  if (!header_token(header, "OBS_OFFSET", tok, sizeof tok) || !parse_number(tok, out->obs_offset)) {
    if (!header_token(header, "OFFSET", tok, sizeof tok) || !parse_number(tok, out->obs_offset))
      fail_key("OBS_OFFSET");
  }
  if (!header_token(header, "INSTRUMENT", out->instrument, sizeof out->instrument)) out->instrument[0] = '\0';
  opt_num("HDR_SIZE", out->hdr_size, uint64_t(0));
  opt_num("RESOLUTION", out->resolution, 1u);
  if (rc != DSPX_OK) return rc;
  if (out->ndim != 1 && out->ndim != 2 && out->ndim != 4) return DSPX_E_STATE;  // ASCIIObservation.C:241-252
  return DSPX_OK;
}

}  // extern "C"
