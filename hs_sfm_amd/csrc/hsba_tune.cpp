// hsba_tune.cpp — the project's autotuner: process-wide table of measured
// kernel choices.
//
// A choice is keyed by a class of windows (precision, camera block width E,
// incidence bucket = floor(log2(mean incidences per camera pair))) and is
// made the first time an engine meets the class: the engine times every
// candidate on that class's own pairs and keeps the fastest.  Because the key
// is a property of each window, a window solved alone and the same window
// inside a batch get the same kernel, so their arithmetic stays identical.
// With HSBA_TUNE_FILE set, the table is read from that file at first use and
// written back after every new decision.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "hsba_host.h"

namespace hsba {

namespace {

struct Entry {
  int variant;
  double ms[kPairVariants];
  int64_t pairs;
};

typedef std::tuple<int, int, int> Key;  // precision, E, bucket

std::mutex table_mu;
std::map<Key, Entry> table;
bool loaded = false;

void load_locked() {
  if (loaded) return;
  loaded = true;
  const char *path = std::getenv("HSBA_TUNE_FILE");
  if (!path) return;
  FILE *f = std::fopen(path, "r");
  if (!f) return;
  int prec, e, b, v;
  double m0, m1, m2;
  long long np;
  while (std::fscanf(f, "%d %d %d %d %lf %lf %lf %lld", &prec, &e, &b, &v, &m0, &m1, &m2, &np) == 8)
    if (v >= 0 && v < kPairVariants) table[Key(prec, e, b)] = Entry{v, {m0, m1, m2}, (int64_t)np};
  std::fclose(f);
}

void save_locked() {
  const char *path = std::getenv("HSBA_TUNE_FILE");
  if (!path) return;
  FILE *f = std::fopen(path, "w");
  if (!f) return;
  for (const auto &kv : table)
    std::fprintf(f, "%d %d %d %d %.6g %.6g %.6g %lld\n", std::get<0>(kv.first), std::get<1>(kv.first),
                 std::get<2>(kv.first), kv.second.variant, kv.second.ms[0], kv.second.ms[1], kv.second.ms[2],
                 (long long)kv.second.pairs);
  std::fclose(f);
}

}  // namespace

std::mutex &tune_mutex() { return table_mu; }

bool tune_lookup_locked(int precision, int E, int bucket, int *variant) {
  load_locked();
  auto it = table.find(Key(precision, E, bucket));
  if (it == table.end()) return false;
  *variant = it->second.variant;
  return true;
}

void tune_store_locked(int precision, int E, int bucket, int variant, const double ms[kPairVariants],
                       int64_t pairs) {
  table[Key(precision, E, bucket)] = Entry{variant, {ms[0], ms[1], ms[2]}, pairs};
  save_locked();
}

}  // namespace hsba

extern "C" {

int hsba_tuning_table(hsba_tune_entry *entries, int32_t max_entries, int32_t *count) {
  return hsba::guarded([&]() -> int {
    if (!count || (max_entries > 0 && !entries)) return HSBA_ERR_INVALID;
    std::lock_guard<std::mutex> g(hsba::table_mu);
    hsba::load_locked();
    int32_t n = 0;
    for (const auto &kv : hsba::table) {
      if (n < max_entries) {
        hsba_tune_entry &o = entries[n];
        o.precision = std::get<0>(kv.first);
        o.camera_block = std::get<1>(kv.first);
        o.bucket = std::get<2>(kv.first);
        o.variant = kv.second.variant;
        for (int v = 0; v < hsba::kPairVariants; ++v) o.ms[v] = kv.second.ms[v];
        o.pairs = kv.second.pairs;
      }
      ++n;
    }
    *count = n;
    return HSBA_OK;
  });
}

int hsba_tuning_reset(void) {
  return hsba::guarded([&]() -> int {
    std::lock_guard<std::mutex> g(hsba::table_mu);
    hsba::table.clear();
    hsba::loaded = true;  // a reset table is not reloaded from the file
    return HSBA_OK;
  });
}

}  // extern "C"
