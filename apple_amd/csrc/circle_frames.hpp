// Byte builders of the Circle FRI transcript V1 frames and the 44-byte domain descriptor
// (CircleProofFormat.swift:71-127 descriptor codec, :906-975 frames), shared by the prover's
// setup and the host gates.
#pragma once
#include <stdint.h>

#include <vector>

namespace azk::frames {

inline void put32(std::vector<uint8_t> &v, uint32_t x) { for (int k = 0; k < 4; ++k) v.push_back((uint8_t)(x >> (8 * k))); }
inline void put64(std::vector<uint8_t> &v, uint64_t x) { for (int k = 0; k < 8; ++k) v.push_back((uint8_t)(x >> (8 * k))); }

// u32 len || "AppleZKProver.CircleFRI.PCS.V1" || u32 1 || u8 type  (:966-973)
inline std::vector<uint8_t> base_frame(uint8_t type)
{
    static const char dom[] = "AppleZKProver.CircleFRI.PCS.V1";
    std::vector<uint8_t> f;
    put32(f, (uint32_t)(sizeof dom - 1));
    f.insert(f.end(), dom, dom + sizeof dom - 1);
    put32(f, 1);
    f.push_back(type);
    return f;
}

// 44-byte canonical domain descriptor (CircleDomainDescriptorCodecV1.encode, :75-93).
inline std::vector<uint8_t> descriptor(uint32_t log_n)
{
    std::vector<uint8_t> d = {0x41, 0x5a, 0x4b, 0x43, 0x44, 0x56, 0x31, 0x00};
    for (uint32_t w : {1u, 0x7fffffffu, 31u, 2u, 1268011823u, log_n, 1u << (31 - (log_n + 1)), log_n - 1, 1u}) put32(d, w);
    return d;
}

// Header, descriptor, security and public-input frames plus the digest: the constant
// transcript prefix (prefixTranscriptAndChallenges, :796-818), as byte messages.
inline std::vector<std::vector<uint8_t>> prefix_messages(uint32_t log_n, uint32_t log_blowup, uint32_t queries,
                                                         uint32_t folding, uint32_t grinding, const uint8_t digest[32])
{
    std::vector<std::vector<uint8_t>> m;
    std::vector<uint8_t> f = base_frame(0);
    for (uint32_t w : {1u, 1u, 0x7fffffffu, 2u, 1268011823u}) put32(f, w);
    m.push_back(f);
    m.push_back(descriptor(log_n));
    f = base_frame(1);
    for (uint32_t w : {log_blowup, queries, folding, grinding, log_blowup * queries + grinding}) put32(f, w);
    m.push_back(f);
    f = base_frame(2);
    put32(f, 32);
    m.push_back(f);
    m.emplace_back(digest, digest + 32);
    return m;
}

inline std::vector<uint8_t> commitment_frame(uint32_t r)
{
    std::vector<uint8_t> c = base_frame(3);
    put32(c, r);
    put32(c, 32);
    return c;
}

inline std::vector<uint8_t> challenge_frame(uint32_t r)
{
    std::vector<uint8_t> c = base_frame(4);
    put32(c, r);
    put32(c, 0x7fffffffu);
    put32(c, 4);
    return c;
}

} // namespace azk::frames
