// Native PlaSim restart files (plasim_restart / plasim_status): Fortran
// unformatted sequential records, each a little-endian uint32 byte count,
// the bytes, and the same count again. put_restart_array writes a field as
// two records: its name (16 characters, blank-padded) and its values (FP32
// or FP64, as the model was built). NATIVE_RESTART.patch
// (src/model/climate/gcm/checkpoint/native/index.ts:3-46) overwrites named
// fields in such a file; the Wasm runner uses it to carry a continuation's
// edits into the CPU engine's restart (runner/wasm/gcm.worker.ts:112-119).
//
// Host byte work over a file of a few MB: nothing here touches the device.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exoplasim_mi355x.h"
#include "engine.h"

namespace {

uint32_t read_u32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);  // little-endian host
  return v;
}

struct Record {
  size_t start, bytes;
};

// One record at `offset`, advancing it; "" or the reference's error text.
std::string next_record(const uint8_t* data, size_t size, size_t& offset, Record& r) {
  if (offset + 8 > size) return "Truncated native restart record";
  const size_t bytes = read_u32(data + offset);
  const size_t start = offset + 4;
  if (start + bytes + 4 > size || read_u32(data + start + bytes) != bytes) return "Invalid native restart record";
  offset = start + bytes + 4;
  r = Record{start, bytes};
  return "";
}

// JavaScript's String.prototype.trim over the ASCII a Fortran name holds.
std::string trimmed(const uint8_t* p, size_t n) {
  auto space = [](uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; };
  size_t a = 0, b = n;
  while (a < b && space(p[a])) a++;
  while (b > a && space(p[b - 1])) b--;
  return std::string(reinterpret_cast<const char*>(p) + a, b - a);
}

// Walks every (name, data) pair; visit(name, data record) returns "" to go
// on or an error text to stop with.
template <class Visit>
std::string walk(const uint8_t* data, size_t size, Visit visit) {
  size_t offset = 0;
  while (offset < size) {
    Record name, values;
    std::string err = next_record(data, size, offset, name);
    if (!err.empty()) return err;
    if (name.bytes != 16) return "Invalid native restart field name";
    const std::string key = trimmed(data + name.start, 16);
    err = next_record(data, size, offset, values);
    if (!err.empty()) return err;
    err = visit(key, values);
    if (!err.empty()) return err;
  }
  return "";
}

}  // namespace

extern "C" int exo_native_find(const void* restart, size_t bytes, const char* name, size_t* offset,
                               size_t* data_bytes) {
  if (!restart || !name || !offset || !data_bytes) return exo::set_error(EXO_E_ARG, "missing argument");
  bool found = false;
  const std::string want(name);
  const std::string err = walk(static_cast<const uint8_t*>(restart), bytes, [&](const std::string& key, Record r) {
    if (!found && key == want) {
      found = true;
      *offset = r.start;
      *data_bytes = r.bytes;
    }
    return std::string();
  });
  if (!err.empty()) return exo::set_error(EXO_E_FORMAT, err);
  if (!found) return exo::set_error(EXO_E_ARG, "Native restart fields missing: " + want);
  return EXO_OK;
}

extern "C" int exo_native_patch(const void* restart, size_t bytes, int n_fields, const char* const* names,
                                const double* const* values, const size_t* counts, void* out) {
  if (!restart || !out || n_fields < 0 || (n_fields > 0 && (!names || !values || !counts)))
    return exo::set_error(EXO_E_ARG, "missing argument");
  uint8_t* result = static_cast<uint8_t*>(out);
  if (out != restart) std::memmove(result, restart, bytes);
  std::vector<bool> pending(n_fields, true);
  const std::string err = walk(result, bytes, [&](const std::string& key, Record r) {
    // a later field of the same name wins, as an object's keys do
    int k = -1;
    for (int i = n_fields - 1; i >= 0; i--)
      if (names[i] && key == names[i]) {
        k = i;
        break;
      }
    if (k < 0) return std::string();
    const size_t n = counts[k];
    const size_t precision = n ? r.bytes / n : 0;
    if (n == 0 || r.bytes % n != 0 || (precision != 4 && precision != 8))
      return "Native restart shape mismatch for " + key;
    for (size_t i = 0; i < n; i++) {
      if (precision == 8) {
        std::memcpy(result + r.start + 8 * i, &values[k][i], 8);
      } else {
        const float f = (float)values[k][i];
        std::memcpy(result + r.start + 4 * i, &f, 4);
      }
    }
    for (int i = 0; i < n_fields; i++)
      if (names[i] && key == names[i]) pending[i] = false;
    return std::string();
  });
  if (!err.empty()) return exo::set_error(EXO_E_FORMAT, err);
  std::string missing;
  std::vector<std::string> listed;
  for (int i = 0; i < n_fields; i++) {
    if (!pending[i]) continue;
    const std::string key = names[i] ? names[i] : "";
    bool dup = false;
    for (const auto& s : listed) dup |= s == key;
    if (dup) continue;
    listed.push_back(key);
    missing += (missing.empty() ? "" : ", ") + key;
  }
  if (!missing.empty()) return exo::set_error(EXO_E_FORMAT, "Native restart fields missing: " + missing);
  return EXO_OK;
}
