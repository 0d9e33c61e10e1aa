// ref_hash.cpp — TEST INFRASTRUCTURE ONLY. A C entry point for the reference's own content hash
// (core/hash/include/core/hash/hash.h, hash_bytes with the engine's seed, which the `.clusters`
// header holds), compiled from the reference's header where it lies (oracle/Makefile, target
// `ref`) into oracle/_ref/. The tests hold the backend's `.clusters` reader to it.
#include <core/hash/hash.h>

#include <cstdint>

extern "C" uint64_t ge_ref_hash_bytes(const void* bytes, uint64_t size) { return engine::hash_bytes(bytes, size); }
