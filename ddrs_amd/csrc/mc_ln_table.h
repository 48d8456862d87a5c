// Generated by gen_ln_table.py: the reduction table of ln_tab (mc_math.h).
// {invc, logc}: invc near 1/c for the middle c of the entry's interval,
// logc = -ln(invc) rounded to double; entry 18 holds 1.0 exactly.
#pragma once

namespace ddrs {

constexpr unsigned kLnOff = 0x3f360000u;  // bits of 0.7109375
constexpr int kLnBits = 5;
constexpr int kLnEntries = 32;

struct LnEnt {
    double invc, logc;
};

DDRS_TABLE LnEnt kLnTab[kLnEntries] = {
    {0x1.642c8590b2164p+0, -0x1.522ae0738a3d7p-2},
    {0x1.5c9882b931057p+0, -0x1.3c25277333183p-2},
    {0x1.5555555555555p+0, -0x1.269621134db91p-2},
    {0x1.4e5e0a72f0539p+0, -0x1.1178e8227e47ap-2},
    {0x1.47ae147ae147bp+0, -0x1.f991c6cb3b37ap-3},
    {0x1.4141414141414p+0, -0x1.d1037f2655e7bp-3},
    {0x1.3b13b13b13b14p+0, -0x1.a93ed3c8ad9e5p-3},
    {0x1.3521cfb2b78c1p+0, -0x1.823c16551a3c0p-3},
    {0x1.2f684bda12f68p+0, -0x1.5bf406b543db0p-3},
    {0x1.29e4129e4129ep+0, -0x1.365fcb0159014p-3},
    {0x1.2492492492492p+0, -0x1.1178e8227e47ap-3},
    {0x1.1f7047dc11f70p+0, -0x1.da7276384469ep-4},
    {0x1.1a7b9611a7b96p+0, -0x1.9335e5d594988p-4},
    {0x1.15b1e5f75270dp+0, -0x1.4d3115d207eacp-4},
    {0x1.1111111111111p+0, -0x1.08598b59e3a06p-4},
    {0x1.0c9714fbcda3bp+0, -0x1.894aa149fb34bp-5},
    {0x1.0842108421084p+0, -0x1.0415d89e74440p-5},
    {0x1.0410410410410p+0, -0x1.0205658935837p-6},
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.f07c1f07c1f08p-1, 0x1.f829b0e7832f8p-6},
    {0x1.e1e1e1e1e1e1ep-1, 0x1.f0a30c01162a8p-5},
    {0x1.d41d41d41d41dp-1, 0x1.6f0d28ae56b4ep-4},
    {0x1.c71c71c71c71cp-1, 0x1.e27076e2af2eap-4},
    {0x1.bacf914c1bad0p-1, 0x1.29552f81ff521p-3},
    {0x1.af286bca1af28p-1, 0x1.5ff3070a793d6p-3},
    {0x1.a41a41a41a41ap-1, 0x1.9525a9cf456b6p-3},
    {0x1.999999999999ap-1, 0x1.c8ff7c79a9a20p-3},
    {0x1.8f9c18f9c18fap-1, 0x1.fb9186d5e3e29p-3},
    {0x1.8618618618618p-1, 0x1.1675cababa60fp-2},
    {0x1.7d05f417d05f4p-1, 0x1.2e8e2bae11d31p-2},
    {0x1.745d1745d1746p-1, 0x1.4618bc21c5ec2p-2},
    {0x1.6c16c16c16c17p-1, 0x1.5d1bdbf5809cap-2},
};

}  // namespace ddrs
