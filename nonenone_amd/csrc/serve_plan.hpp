// serve_plan.hpp — merge plan of one serving batch (host-only, no HIP types).
//
// Steps whose phase updates may run as ONE launch over their gathered cells are
// those that share a fluid (the caller's fluid_group, >= 0) AND the cell-state
// layout that the gather and scatter copy: primary-vector length np, phase-state
// length nps and the well-unknown flag.  Two steps of the same fluid whose
// layouts differ (one with a well unknown, one without) form separate groups;
// merging them would copy ncell * np_lead doubles into a member buffer sized
// ncell * np_member and decode the cells with the wrong layout.
#ifndef NN_SERVE_PLAN_HPP
#define NN_SERVE_PLAN_HPP

#include <cstddef>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

namespace nn
{

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Staging of a merged phase update: T, primary, phase state, status, flags.
inline size_t stagingBytes(int64_t cells, int np, int nps)
{
    return align256(size_t(cells) * sizeof(double)) +
           align256(size_t(cells) * np * sizeof(double)) +
           align256(size_t(cells) * nps * sizeof(double)) +
           align256(size_t(cells) * sizeof(int32_t)) + align256(size_t(cells));
}

// Merge groups of a batch in first-appearance order; fluid_group < 0 stands alone.
struct GroupPlan
{
    using Key = std::tuple<int64_t, int, int, int>;  // fluid group, np, nps, well flag
    std::vector<std::vector<int64_t>> groups;
    std::map<Key, size_t> index;
    std::vector<int64_t> cells;
    std::vector<int> np, nps, well;

    void add(int64_t ticket, int64_t fluidGroup, int64_t ncell, int np_, int nps_, int well_)
    {
        if (fluidGroup >= 0)
        {
            const Key key{fluidGroup, np_, nps_, well_ != 0 ? 1 : 0};
            auto it = index.find(key);
            if (it != index.end())
            {
                groups[it->second].push_back(ticket);
                cells[it->second] += ncell;
                return;
            }
            index[key] = groups.size();
        }
        groups.push_back({ticket});
        cells.push_back(ncell);
        np.push_back(np_);
        nps.push_back(nps_);
        well.push_back(well_ != 0 ? 1 : 0);
    }

    size_t staging() const
    {
        size_t b = 0;
        for (size_t g = 0; g < groups.size(); ++g)
            if (groups[g].size() > 1)
                b += stagingBytes(cells[g], np[g], nps[g]);
        return b;
    }
};

} // namespace nn

#endif
