// Host build of the device header's table-driven exp and sincos, for tests/test_device_math.py: reads
// arguments from stdin ("e a" or "t y" per line) and prints e^{-a}, or sin y and cos y, as hex doubles.
#include "../surfaceonlydeformables_amd/csrc/cqbem_device.hpp"

#include <cstdio>

int main() {
    static const double E[64][2] = CQBEM_EXP_TABLE;
    static const double T[64][4] = CQBEM_TRIG_TABLE;
    cqbem::MathTables M;
    for (int k = 0; k < 64; k++) {
        M.e[k] = make_double2(E[k][0], E[k][1]);
        M.t[k] = make_double4(T[k][0], T[k][1], T[k][2], T[k][3]);
    }
    char kind;
    double x;
    while (std::scanf(" %c %lf", &kind, &x) == 2) {
        if (kind == 'e') {
            std::printf("%a\n", cqbem::exp_neg_tab(x, M.e));
        } else {
            double s, c;
            cqbem::sincos_tab(x, M.t, &s, &c);
            std::printf("%a %a\n", s, c);
        }
    }
    return 0;
}
