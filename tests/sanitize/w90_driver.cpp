// Drives the Wannier90 reader and operator pipeline (sbewh_amd/csrc/w90.cpp)
// over the reference's example files, for the AddressSanitizer /
// UndefinedBehaviorSanitizer build of tests/test_sanitizers.py.
#include <cstdio>
#include <string>

#include "sbewh_w90.h"

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string seed = argv[1];
    sbe_w90_tb *tb = nullptr;
    sbe_w90_wsvec *ws = nullptr;
    if (sbe_w90_tb_read((seed + "_tb.dat").c_str(), &tb)) return 3;
    if (sbe_w90_wsvec_read((seed + "_wsvec.dat").c_str(), &ws)) return 4;
    for (int flags = 0; flags < 4; flags++)
        for (int dim = 0; dim <= 3; dim++) {
            sbe_w90_model *m = nullptr;
            if (sbe_w90_model_build(tb, flags & 1 ? ws : nullptr, flags, dim, &m)) return 5;
            sbe_model v{};
            sbe_w90_model_view(m, &v);
            sbe_w90_model_free(m);
        }
    int r[3], cnt = 0, T[64];
    for (int i = 0; i < sbe_w90_wsvec_count(ws); i++) sbe_w90_wsvec_entry(ws, i, r, nullptr, nullptr, &cnt, T, 16);
    sbe_w90_tb *bad = nullptr;
    if (sbe_w90_tb_read("/nonexistent_tb.dat", &bad) == 0) return 6;
    sbe_w90_wsvec_free(ws);
    sbe_w90_tb_free(tb);
    std::printf("w90 ok\n");
    return 0;
}
