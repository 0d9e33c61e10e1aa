// The `.clusters` reader under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU (test
// infrastructure; tests/test_sanitized.py builds and runs it):
//
//   clusters_driver fuzz <file> <iterations> <seed>
//     Mutates a valid container (byte and word overwrites, boundary values in the header and the
//     section table, truncation, growth) and, for three mutants in four, repairs the content hash
//     so the reader gets past it to the sections. Every mutant must be refused with
//     GE_RT_ERROR_INVALID_SCENE and a message, or read; a read scene is walked end to end (every
//     cluster's triangles and vertices in range of its streams) and freed. Any out-of-bounds
//     access, use after free, leak or undefined operation aborts the run with the sanitizer's
//     report.
//   clusters_driver selfcheck
//     Reads one byte past a heap block, which the sanitizer must report.
//   clusters_driver oom <file>
//     Fails the program's n-th host allocation, for n = 1, 2, ... until a read completes: each
//     failure must come back as GE_RT_ERROR_OUT_OF_MEMORY with no scene and the message "out of
//     host memory" (and free whatever was built; the leak check sees to it); then the read must
//     succeed.
//
// The reader is compiled into this program from its own source.
#include "../../game_engine_amd/csrc/rt_clusters.hip"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <map>

// ---- the allocation hook: the program's global operator new, failing on request --------------
static long g_fail_at = 0;  // 0: never; n: the n-th allocation from arming fails
static long g_count = 0;

static bool injected_failure() {
  if (g_fail_at == 0) return false;
  return ++g_count == g_fail_at;
}
void* operator new(std::size_t n) {
  if (injected_failure()) throw std::bad_alloc();
  if (void* p = std::malloc(n ? n : 1)) return p;
  throw std::bad_alloc();
}
void* operator new[](std::size_t n) { return operator new(n); }
void* operator new(std::size_t n, const std::nothrow_t&) noexcept {
  if (injected_failure()) return nullptr;
  return std::malloc(n ? n : 1);
}
void* operator new[](std::size_t n, const std::nothrow_t& t) noexcept { return operator new(n, t); }
void operator delete(void* p) noexcept { std::free(p); }
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete(void* p, std::size_t) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

static void arm(long n) {
  g_count = 0;
  g_fail_at = n;
}

// ---- helpers --------------------------------------------------------------------------------------
static std::vector<uint8_t> load(const char* path) {
  std::ifstream f(path, std::ios::binary);
  return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

struct Rng {
  uint64_t s;
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

static void rehash(std::vector<uint8_t>& b) {
  if (b.size() < 32) return;
  uint64_t total;
  std::memcpy(&total, b.data() + 8, 8);
  if (total < 32 || total > b.size()) return;
  const uint64_t h = ge_clusters_hash(b.data() + 32, total - 32);
  std::memcpy(b.data() + 24, &h, 8);
}

// Touches everything a read scene hands out, as an uploader would.
static uint64_t walk(const ge_clusters* c) {
  ge_scene_desc d;
  ge_clusters_scene(c, &d);
  uint32_t all = 0, levels = 0;
  ge_clusters_counts(c, &all, &levels);
  uint64_t sum = all + levels;
  const ge_mesh_desc& mesh = d.meshes[0];
  for (uint32_t k = 0; k < mesh.cluster_count; ++k) {
    const ge_cluster_desc& cl = d.clusters[mesh.first_cluster + k];
    for (uint32_t t = 0; t < cl.triangle_count; ++t) {
      const uint32_t p = d.triangles[cl.triangle_offset + t];
      for (int j = 0; j < 3; ++j) {
        const uint32_t v = cl.vertex_offset + ((p >> (8 * j)) & 0xffu);
        const uint16_t* q = d.quantized + size_t(v) * 3;
        sum += q[0] + q[1] + q[2] + (d.attributes ? d.attributes[v].normal_oct : 0u);
      }
    }
    sum += d.cluster_materials[mesh.first_cluster + k];
  }
  for (uint32_t m = 0; m < d.material_count; ++m) sum += d.materials[m].flags;
  sum += d.instances[0].mesh;
  return sum;
}

static int fuzz(const std::vector<uint8_t>& seed, long iterations, uint64_t rng_seed) {
  Rng rng{rng_seed * 0x9E3779B97F4A7C15ull + 1};
  std::map<int32_t, long> outcomes;
  uint64_t touched = 0;
  char error[256];
  for (long it = 0; it < iterations; ++it) {
    std::vector<uint8_t> b = seed;
    const int edits = 1 + int(rng.below(4));
    for (int e = 0; e < edits; ++e) {
      const uint64_t kind = rng.below(8);
      if (kind <= 1 && !b.empty()) {  // a byte
        b[rng.below(b.size())] = uint8_t(rng.next());
      } else if (kind == 2 && b.size() >= 4) {  // a word anywhere
        const uint32_t w = uint32_t(rng.next());
        std::memcpy(b.data() + rng.below(b.size() - 3), &w, 4);
      } else if (kind == 3 && b.size() >= 32) {  // a boundary value into the header or section table
        static const uint64_t edges[] = {0, 1, 2, 3, 0xff, 0xffff, 0x7fffffff, 0xffffffff, 0x100000000ull,
                                         0xffffffffffffffffull, 32, 56};
        uint32_t table;
        std::memcpy(&table, b.data() + 12, 4);
        const uint64_t span = std::min<uint64_t>(b.size(), 32 + uint64_t(std::min<uint32_t>(table, 64)) * 24);
        const uint64_t at = rng.below(span / 4) * 4;
        const uint64_t v = edges[rng.below(sizeof(edges) / sizeof(edges[0]))];
        std::memcpy(b.data() + at, &v, std::min<uint64_t>(rng.below(2) ? 8 : 4, b.size() - at));
      } else if (kind == 4) {  // truncation
        b.resize(rng.below(b.size() + 1));
      } else if (kind == 5) {  // growth, the header's length following in half the cases
        const size_t extra = size_t(rng.below(256));
        for (size_t i = 0; i < extra; ++i) b.push_back(uint8_t(rng.next()));
        if (rng.below(2) && b.size() >= 32) {
          const uint64_t total = b.size();
          std::memcpy(b.data() + 8, &total, 8);
        }
      } else if (kind == 6 && b.size() >= 32) {  // a section's count or offset nudged
        uint32_t table;
        std::memcpy(&table, b.data() + 12, 4);
        if (table && 32 + uint64_t(table) * 24 <= b.size()) {
          uint8_t* s = b.data() + 32 + rng.below(table) * 24;
          uint64_t v;
          const int field = 8 + 8 * int(rng.below(2));
          std::memcpy(&v, s + field, 8);
          v += uint64_t(int64_t(rng.below(9)) - 4);
          std::memcpy(s + field, &v, 8);
        }
      } else if (!b.empty()) {  // a run of one byte value
        const size_t at = size_t(rng.below(b.size())), n = std::min<size_t>(b.size() - at, rng.below(64));
        std::memset(b.data() + at, int(rng.below(2) ? 0 : 0xff), n);
      }
    }
    if (rng.below(4) != 0) rehash(b);
    ge_clusters* c = nullptr;
    error[0] = '\0';
    const int32_t r = ge_clusters_read(b.empty() ? nullptr : b.data(), b.size(), uint32_t(rng.below(2)), &c,
                                       error, sizeof(error));
    ++outcomes[r];
    if (r == GE_RT_OK) {
      if (!c) return std::fprintf(stderr, "iteration %ld: read with no scene\n", it), 1;
      touched += walk(c);
      ge_clusters_destroy(c);
    } else {
      if (c) return std::fprintf(stderr, "iteration %ld: refused yet a scene\n", it), 1;
      if (r != GE_RT_ERROR_INVALID_SCENE || !error[0])
        return std::fprintf(stderr, "iteration %ld: status %d, message '%s'\n", it, r, error), 1;
    }
  }
  std::printf("fuzz: %ld mutants:", iterations);
  for (const auto& [code, n] : outcomes) std::printf(" status %d x %ld,", code, n);
  std::printf(" checksum %llu\n", static_cast<unsigned long long>(touched));
  return outcomes[GE_RT_OK] > 0 && outcomes[GE_RT_ERROR_INVALID_SCENE] > 0 ? 0 : 1;
}

static int oom(const std::vector<uint8_t>& file) {
  char error[256];
  for (long n = 1; n < 100000; ++n) {
    ge_clusters* c = nullptr;
    arm(n);
    const int32_t r = ge_clusters_read(file.data(), file.size(), GE_CLUSTERS_IMAGES, &c, error, sizeof(error));
    arm(0);
    if (r == GE_RT_OK) {
      if (!c) return std::fprintf(stderr, "no scene after a read\n"), 1;
      walk(c);
      ge_clusters_destroy(c);
      std::printf("oom: %ld allocation points failed one at a time, each GE_RT_ERROR_OUT_OF_MEMORY; then read\n",
                  n - 1);
      return n > 2 ? 0 : 1;
    }
    if (r != GE_RT_ERROR_OUT_OF_MEMORY || c || std::strcmp(error, "out of host memory") != 0)
      return std::fprintf(stderr, "allocation %ld: status %d, scene %p, message '%s'\n", n, r, static_cast<void*>(c),
                          error),
             1;
  }
  return std::fprintf(stderr, "the read never completed\n"), 1;
}

// One read a byte past a heap block: the sanitizer must stop the program here (its own check that
// the build is instrumented).
static int selfcheck() {
  std::vector<uint8_t> block(8, 1);
  const volatile uint8_t* p = block.data();
  std::printf("read %u\n", unsigned(p[block.size()]));
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "selfcheck") == 0) return selfcheck();
  if (argc >= 5 && std::strcmp(argv[1], "fuzz") == 0) {
    const std::vector<uint8_t> seed = load(argv[2]);
    if (seed.empty()) return std::fprintf(stderr, "cannot read %s\n", argv[2]), 2;
    return fuzz(seed, std::atol(argv[3]), std::strtoull(argv[4], nullptr, 10));
  }
  if (argc >= 3 && std::strcmp(argv[1], "oom") == 0) {
    const std::vector<uint8_t> file = load(argv[2]);
    if (file.empty()) return std::fprintf(stderr, "cannot read %s\n", argv[2]), 2;
    return oom(file);
  }
  std::fprintf(stderr, "usage: clusters_driver fuzz <file> <iterations> <seed> | oom <file>\n");
  return 2;
}
