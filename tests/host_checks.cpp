// Host checks of the library's device-free decisions, built and run by tests/test_host_checks.py
// with g++; prints "ok" or the failed check:
//   LaunchLedger (csrc/ba_ledger.hpp): the event-pair bookkeeping of a profiled kernel;
//   use_explicit_product (csrc/ba_policy.hpp): the PCG product form of an LM iteration.
#include <cstdio>
#include <cstdlib>

#include "../openmvg-rs_amd/csrc/ba_ledger.hpp"
#include "../openmvg-rs_amd/csrc/ba_policy.hpp"

namespace {
int created = 0, destroyed = 0;
struct CountingOps {
    void create(int& e) { e = ++created; }
    void destroy(int) { ++destroyed; }
};
#define CHECK(cond)                                                  \
    do {                                                             \
        if (!(cond)) {                                               \
            std::printf("failed: %s (line %d)\n", #cond, __LINE__);  \
            std::exit(1);                                            \
        }                                                            \
    } while (0)
}  // namespace

void check_product_policy() {
    using ba::use_explicit_product;
    // AUTO: the explicit system after a solve of 10 or more PCG iterations, not after 9
    CHECK(!use_explicit_product(true, BA_SCHUR_AUTO, 9));
    CHECK(use_explicit_product(true, BA_SCHUR_AUTO, 10));
    CHECK(use_explicit_product(true, BA_SCHUR_AUTO, 600));
    // before the first solve, and after a solve whose preconditioner failed (0 iterations)
    CHECK(!use_explicit_product(true, BA_SCHUR_AUTO, 0));
    // forced forms
    CHECK(use_explicit_product(true, BA_SCHUR_EXPLICIT, 0));
    CHECK(!use_explicit_product(true, BA_SCHUR_IMPLICIT, 600));
    // no explicit system set up (another topology, or it did not fit): never
    CHECK(!use_explicit_product(false, BA_SCHUR_EXPLICIT, 600));
    CHECK(!use_explicit_product(false, BA_SCHUR_AUTO, 600));
}

int main() {
    check_product_policy();
    {
        ba::LaunchLedger<int, CountingOps> L;
        CHECK(L.count() == 0 && L.capacity() == 0);
        // a batch of four launches
        for (int k = 0; k < 4; ++k) L.next();
        CHECK(L.count() == 4 && L.capacity() == 4 && created == 8);
        // the device reports more products than launches were recorded (a stale or corrupt
        // counter): the count is clamped to the pairs that exist
        L.keep(4 + 7);
        CHECK(L.count() == 4);
        // ... and a later launch gets a new pair instead of reading past the vector
        auto& e = L.next();
        CHECK(L.count() == 5 && L.capacity() == 5 && e.first == 9 && e.second == 10);
        // two of the next batch's four launches applied a product: the rest are reused
        const size_t before = L.count();
        for (int k = 0; k < 4; ++k) L.next();
        CHECK(L.count() == 9 && L.capacity() == 9);
        L.keep(before + 2);
        CHECK(L.count() == 7 && L.capacity() == 9);
        const int reused = L.next().first;
        CHECK(reused == L[7].first && L.count() == 8 && created == 18);
        // a negative difference is never passed (the caller clamps at zero); keep(0) empties
        L.keep(0);
        CHECK(L.count() == 0 && L.capacity() == 9);
        for (size_t k = 0; k < 12; ++k) L.next();
        CHECK(L.count() == 12 && L.capacity() == 12);
    }
    CHECK(destroyed == created && created == 24);
    std::printf("ok\n");
    return 0;
}
