// Behaviour of the standalone ParamsCarrier against the reference's
// (Input/include/ParamsCarrier.hpp:158-300): non-string puts keep the first
// value, string puts replace, typed gets fall back to the default on a
// missing key or another type, eraseAllItems keeps only the listed keys.
#include <cstdio>
#include <cstdlib>

#include "ParamsCarrier.hpp"

#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                             \
        }                                                             \
    } while (0)

int main() {
    ParamsCarrier *pc = ParamsCarrier::instance();
    CHECK(pc == ParamsCarrier::instance());
    pc->putFloat("dt", 500.0f);
    pc->putFloat("dt", 50.0f);            // kept: first value wins
    CHECK(pc->getFloat("dt", -1.0f) == 500.0f);
    pc->putInt("millions", 2);
    pc->putInt("millions", 7);
    CHECK(pc->getInt("millions", 1) == 2);
    pc->putDouble("x", 1.5);
    pc->putDouble("x", 2.5);
    CHECK(pc->getDouble("x", 0.0) == 1.5);
    pc->putString("model", "1D Bp");
    pc->putString("model", "1D Fp");      // strings are replaced
    CHECK(pc->getString("model", "") == "1D Fp");
    CHECK(pc->getInt("dt", 42) == 42);    // wrong type -> default
    CHECK(pc->getFloat("missing", 3.0f) == 3.0f);
    CHECK(pc->getString("dt", "d") == "d");
    pc->eraseAllItems({"model", "millions"});
    CHECK(pc->getString("model", "") == "1D Fp");
    CHECK(pc->getInt("millions", 1) == 2);
    CHECK(pc->getFloat("dt", -1.0f) == -1.0f);
    CHECK(pc->getDouble("x", 0.0) == 0.0);
    pc->putFloat("dt", 5.0f);             // after erase a new value goes in
    CHECK(pc->getFloat("dt", -1.0f) == 5.0f);
    std::puts("ok");
    return 0;
}
