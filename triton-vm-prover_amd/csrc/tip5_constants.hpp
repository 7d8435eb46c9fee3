// Tip5 specification constants shared by the device kernels and the host
// drop-in classes (plain C++, no HIP).
#pragma once

#include <stdint.h>

namespace tip5c {

constexpr int kStateSize = 16;
constexpr int kRounds = 5;

// Round constants of the Tip5 specification (src/hash/tip5.cpp:48-79).
constexpr uint64_t kRoundConstants[kRounds * kStateSize] = {
    0xbd2a3deb61ab60deull, 0xea7df21ad9547ed2ull, 0x900b3677a1de063full, 0x1b46887e876c8677ull,
    0xd364d977889cfb97ull, 0xdc8dfac843699f02ull, 0x375c405d7190db58ull, 0x27924006d2b0d4b1ull,
    0x78dd1172d483cd38ull, 0x3346c66244882a56ull, 0xb0249b279f498aa5ull, 0x94cd51be79338d4dull,
    0xb0e0dc7052c5b218ull, 0xf8dcc4d248adad95ull, 0x68e3c635fec868b7ull, 0xd7d06b3ffb6b0d8cull,
    0xf3500dea20ef032aull, 0x4865bf175bba5803ull, 0xd5f7fe3027287a27ull, 0xa57333f44e193412ull,
    0x8726e153a977eae2ull, 0x3014a98463fc191bull, 0xba145461af39b212ull, 0x03ab70105933202full,
    0x3d90b7eebfcf71e5ull, 0x386322b1cc520bfdull, 0x27c2c8daf774f675ull, 0x4fcb83f50309bc6aull,
    0x5e6d5ce8275f3cb3ull, 0xecc2f6592c8f905cull, 0x837f532461e609b4ull, 0xb2b1f6b95c92c93cull,
    0xc0027af556411dc1ull, 0x16e18c885fc2a26cull, 0x8880ef183d9f2bf3ull, 0xb2930bdb5ca88c45ull,
    0x9c2ec8322e1c1553ull, 0xe5b05eaf3220a674ull, 0xa49cc6ae4b861c4eull, 0x11708e0aeb86ebd7ull,
    0xc09de92bbc3902e0ull, 0x929b3c79516bcbc1ull, 0xe006e5bf738f27d1ull, 0x2d9e1ec0eac8ea38ull,
    0x0984d8d94bf937c5ull, 0x4959273c220e6747ull, 0xfe1d934207e796faull, 0x2b9b9298f2f6dd73ull,
    0x07a1f5a67d6e3a41ull, 0x4407593ee73743d9ull, 0x9f054720ef802e59ull, 0x78d4b711336e6aa6ull,
    0xadc638aef3c8b228ull, 0xa4d6d3e86afb2114ull, 0x9d4808e725531968ull, 0x369804df3866d0efull,
    0xe6dbd9a9d2215024ull, 0x8ed22ca212ee85b2ull, 0x397bb882fcd23eb6ull, 0xeb8f8786d7277531ull,
    0x9999d4cdaff543b5ull, 0xf382a61217f192d6ull, 0x49c37260b026adc1ull, 0x3ff8918ce35c1019ull,
    0x2e7df8b76080bd07ull, 0xf5dbac250b8a28b9ull, 0x853c3727ae9da4ccull, 0xb2f1f5f3d9e5a26dull,
    0x3fce22012d337847ull, 0x6b5a3e6db7eee347ull, 0x171582cd59dde50dull, 0xc0c0b3095ee62a8aull,
    0x665b25c6f6a203d2ull, 0x3099aed93b6ae69full, 0x801df6092be69c38ull, 0x8066ad0cdfff43cdull,
    0x8af9d44a5f4fdc6bull, 0xd80219cd97c0d762ull, 0x10c9ceba14148ebbull, 0x539bd4c3f2f24474ull,
};

// First column of the circulant MDS matrix: SHA-256("Tip5") read as 16
// little-endian u16 (src/hash/tip5.cpp:27-30 holds the same values).
constexpr int64_t kMdsFirstColumn[kStateSize] = {61402, 1108,  28750, 33823, 7454,  43244,
                                                 53865, 12034, 56951, 27521, 41351, 40901,
                                                 12021, 59689, 26798, 17845};

}  // namespace tip5c
