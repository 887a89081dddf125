// Host build of the engine's record codec (fdtdx_amd/csrc/record_codec.h) for the CPU tests:
// reads float32 values from argv[2], writes for format argv[1] the encoded codes (as uint32),
// the decoded round trip and prev + f * (next - prev) interpolations of consecutive value pairs
// with the factor argv[3], in that order, to argv[4].
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../fdtdx_amd/csrc/record_codec.h"

using namespace fdtdx;

static uint32_t enc(int fmt, float v) {
  switch (fmt) {
    case FDTDX_REC_BF16: return enc_small<8, 7, 127, 0>(v);
    case FDTDX_REC_F16: return enc_small<5, 10, 15, 0>(v);
    case FDTDX_REC_E4M3FN: return enc_small<4, 3, 7, 1>(v);
    case FDTDX_REC_E4M3FNUZ: return enc_small<4, 3, 8, 2>(v);
    case FDTDX_REC_E5M2: return enc_small<5, 2, 15, 0>(v);
    case FDTDX_REC_E5M2FNUZ: return enc_small<5, 2, 16, 2>(v);
    default: return f2u(v);
  }
}

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  const int fmt = atoi(argv[1]);
  const float f = (float)atof(argv[3]);
  FILE* in = fopen(argv[2], "rb");
  if (!in) return 3;
  std::vector<float> x;
  float v;
  while (fread(&v, 4, 1, in) == 1) x.push_back(v);
  fclose(in);
  const size_t n = x.size();
  std::vector<uint32_t> codes(n), round(n), interp(n / 2);
  for (size_t i = 0; i < n; ++i) {
    codes[i] = enc(fmt, x[i]);
    round[i] = f2u(rec_round(x[i], fmt));
  }
  for (size_t i = 0; i + 1 < n; i += 2) {
    const float a = rec_round(x[i], fmt), b = rec_round(x[i + 1], fmt);
    interp[i / 2] = f2u(rec_interp(a, b, f, fmt));
  }
  FILE* out = fopen(argv[4], "wb");
  if (!out) return 4;
  fwrite(codes.data(), 4, n, out);
  fwrite(round.data(), 4, n, out);
  fwrite(interp.data(), 4, n / 2, out);
  fclose(out);
  return 0;
}
