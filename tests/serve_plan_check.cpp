// Host check of the serving batch's merge plan (nonenone_amd/csrc/serve_plan.hpp):
// steps of one fluid group merge only when their cell-state layouts agree.
// Prints "ok" and exits 0 when every assertion holds.
#include <cstdio>
#include <cstdlib>

#include "../nonenone_amd/csrc/serve_plan.hpp"

#define CHECK(c)                                                            \
    do                                                                      \
    {                                                                       \
        if (!(c))                                                           \
        {                                                                   \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                       \
        }                                                                   \
    } while (0)

int main()
{
    using nn::GroupPlan;
    // nc = 3: np 10 without a well unknown, 11 with; nps 17 either way.
    GroupPlan p;
    p.add(1, 0, 100, 10, 17, 0);
    p.add(2, 0, 50, 11, 17, 1);   // same fluid, well layout: its own group
    p.add(3, 0, 30, 10, 17, 0);   // merges with ticket 1
    p.add(4, 0, 20, 11, 17, 1);   // merges with ticket 2
    p.add(5, -1, 40, 10, 17, 0);  // never merges
    p.add(6, 1, 10, 19, 26, 0);   // nc = 6 fluid
    CHECK(p.groups.size() == 4);
    CHECK(p.groups[0].size() == 2 && p.groups[0][0] == 1 && p.groups[0][1] == 3);
    CHECK(p.groups[1].size() == 2 && p.groups[1][0] == 2 && p.groups[1][1] == 4);
    CHECK(p.groups[2].size() == 1 && p.groups[2][0] == 5);
    CHECK(p.groups[3].size() == 1 && p.groups[3][0] == 6);
    CHECK(p.cells[0] == 130 && p.cells[1] == 70);
    CHECK(p.np[0] == 10 && p.np[1] == 11 && p.well[0] == 0 && p.well[1] == 1);
    // staging only for merged groups, each sized with its own layout
    CHECK(p.staging() == nn::stagingBytes(130, 10, 17) + nn::stagingBytes(70, 11, 17));
    // a fluid group whose members have different component counts cannot share
    // a layout either
    GroupPlan q;
    q.add(1, 7, 10, 10, 17, 0);
    q.add(2, 7, 10, 19, 26, 0);
    CHECK(q.groups.size() == 2 && q.staging() == 0);
    std::puts("ok");
    return 0;
}
