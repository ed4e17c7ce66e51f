// CountedState container codec (see counted_state.h). Lengths are checked
// before any allocation; every field is a little-endian 64-bit word.
#include "counted_state.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <stdexcept>

namespace faco {
namespace {

// "GFACOST2" and the earlier "GFACOST1", whose last 8 bytes are skipped
constexpr uint64_t kMagic = 0x3254534f43414647ULL;
constexpr uint64_t kLegacyMagic = 0x3154534f43414647ULL;

void require(bool ok, const char* message) {
    if (!ok) throw std::invalid_argument(message);
}

void put(std::vector<uint8_t>& out, uint64_t v) {
    for (int i = 0; i < 8; ++i) out.push_back(static_cast<uint8_t>(v >> (8 * i)));
}

void put_real(std::vector<uint8_t>& out, double v) {
    uint64_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    put(out, bits);
}

struct Reader {
    const uint8_t* data;
    size_t cursor = 0, end;
    Reader(const uint8_t* d, size_t size) : data(d), end(size) {}
    uint64_t word() {
        require(cursor <= end && end - cursor >= 8, "snapshot metadata truncated");
        uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v |= static_cast<uint64_t>(data[cursor++]) << (8 * i);
        return v;
    }
    uint32_t u32() {
        const uint64_t v = word();
        require(v <= UINT32_MAX, "snapshot integer overflow");
        return static_cast<uint32_t>(v);
    }
    double real() {
        const uint64_t bits = word();
        double v;
        std::memcpy(&v, &bits, sizeof v);
        return v;
    }
    std::vector<uint8_t> bytes(uint64_t size) {
        require(cursor <= end && size <= end - cursor, "snapshot buffer truncated");
        std::vector<uint8_t> out(data + cursor, data + cursor + size);
        cursor += static_cast<size_t>(size);
        return out;
    }
};

}  // namespace

bool same_settings(const faco_settings& a, const faco_settings& b) {
    return a.ants == b.ants && a.primary_width == b.primary_width && a.backup_width == b.backup_width &&
           a.ls_width == b.ls_width && a.beta == b.beta && a.retention == b.retention && a.p_best == b.p_best &&
           a.epoch_source_probability == b.epoch_source_probability &&
           a.ls_evaluation_limit == b.ls_evaluation_limit &&
           a.initial_ls_evaluation_limit == b.initial_ls_evaluation_limit;
}

void CountedState::validate() const {
    const faco_settings& c = settings;
    require(dimension >= 3 && colonies >= 1 && colonies <= 128 && c.ants >= 1 && c.ants <= 4096 &&
            tasks.size() == colonies && incumbents.size() == colonies, "snapshot shape or tasks invalid");
    for (const auto& item : incumbents) {
        require(item.tour.size() == dimension && std::isfinite(item.cost) && item.cost > 0,
                "snapshot host incumbent invalid");
        std::vector<uint32_t> nodes = item.tour;
        std::sort(nodes.begin(), nodes.end());
        for (uint32_t i = 0; i < dimension; ++i) require(nodes[i] == i, "snapshot host tour is not a permutation");
    }
    require(progress_evaluation_limit > 0 && progress_evaluation_limit % c.ants == 0 &&
            progress_evaluation_limit / c.ants <= UINT32_MAX &&
            completed_batches <= progress_evaluation_limit / c.ants && (experiment_mask & 0xffffu),
            "snapshot batches, progress budget or mask invalid");
    require(!buffers.empty() && buffers.size() <= 64, "snapshot buffer set invalid");
    std::set<std::string> names;
    for (const auto& b : buffers)
        require(!b.name.empty() && b.name.size() <= 128 && names.insert(b.name).second && b.element_bytes > 0 &&
                b.element_bytes <= 65536 && b.planes >= 1 && b.planes <= 12 && !b.bytes.empty() &&
                b.bytes.size() % (b.element_bytes * b.planes * colonies) == 0,
                "snapshot buffer name, ABI or plane layout invalid");
}

std::vector<uint8_t> CountedState::serialize() const {
    std::vector<uint8_t> out;
    size_t capacity = 256;
    for (const auto& b : buffers) capacity += b.bytes.size() + b.name.size() + 32;
    out.reserve(capacity);
    put(out, kMagic);
    put(out, dimension);
    put(out, colonies);
    const faco_settings& c = settings;
    for (uint64_t v : {uint64_t(c.ants), uint64_t(c.primary_width), uint64_t(c.backup_width), uint64_t(c.ls_width)})
        put(out, v);
    for (double v : {c.beta, c.retention, c.p_best, c.epoch_source_probability}) put_real(out, v);
    put(out, c.ls_evaluation_limit);
    put(out, c.initial_ls_evaluation_limit);
    put(out, experiment_mask);
    put(out, completed_batches);
    put(out, progress_evaluation_limit);
    put(out, tasks.size());
    for (const auto& t : tasks) { put(out, t.first); put(out, t.second); }
    put(out, incumbents.size());
    for (const auto& item : incumbents) {
        put_real(out, item.cost);
        put(out, item.tour.size());
        for (uint32_t v : item.tour) put(out, v);
    }
    put(out, buffers.size());
    for (const auto& b : buffers) {
        put(out, b.name.size());
        out.insert(out.end(), b.name.begin(), b.name.end());
        put(out, b.element_bytes);
        put(out, b.planes);
        put(out, b.bytes.size());
        out.insert(out.end(), b.bytes.begin(), b.bytes.end());
    }
    return out;
}

size_t CountedState::serialized_size() const {
    size_t size = 8 * (19 + 2 * tasks.size());   // 16 header words, three counts, the tasks
    for (const auto& item : incumbents) size += 8 * (2 + item.tour.size());
    for (const auto& b : buffers) size += 8 * 4 + b.name.size() + b.bytes.size();
    return size;
}

CountedState CountedState::deserialize(const uint8_t* data, size_t size) {
    require(data != nullptr && size >= 16, "snapshot file truncated");
    Reader r(data, size);
    const uint64_t magic = r.word();
    require(magic == kMagic || magic == kLegacyMagic, "unknown snapshot format");
    if (magic == kLegacyMagic) r.end -= 8;   // trailing metadata of the older format is skipped
    CountedState s;
    s.dimension = r.u32();
    s.colonies = r.u32();
    faco_settings& c = s.settings;
    c.ants = r.u32(); c.primary_width = r.u32(); c.backup_width = r.u32(); c.ls_width = r.u32();
    c.beta = r.real(); c.retention = r.real(); c.p_best = r.real(); c.epoch_source_probability = r.real();
    c.ls_evaluation_limit = r.word();
    c.initial_ls_evaluation_limit = r.word();
    s.experiment_mask = r.u32();
    s.completed_batches = r.word();
    s.progress_evaluation_limit = r.word();
    const uint64_t tasks = r.word();
    require(tasks <= 128, "snapshot task count out of range");
    for (uint64_t i = 0; i < tasks; ++i) {
        const uint64_t key = r.word(), seed = r.word();
        s.tasks.emplace_back(key, seed);
    }
    const uint64_t incumbents = r.word();
    require(incumbents <= 128, "snapshot incumbent count out of range");
    for (uint64_t i = 0; i < incumbents; ++i) {
        StateIncumbent item;
        item.cost = r.real();
        const uint64_t length = r.word();
        require(length == s.dimension && length <= (r.end - r.cursor) / 8, "snapshot host tour shape or bytes invalid");
        for (uint64_t j = 0; j < length; ++j) item.tour.push_back(r.u32());
        s.incumbents.push_back(std::move(item));
    }
    const uint64_t buffers = r.word();
    require(buffers <= 64, "snapshot buffer count out of range");
    for (uint64_t i = 0; i < buffers; ++i) {
        const uint64_t length = r.word();
        require(length <= 128, "snapshot buffer name out of range");
        const auto name = r.bytes(length);
        StateBuffer b;
        b.name.assign(name.begin(), name.end());
        b.element_bytes = r.word();
        b.planes = r.word();
        b.bytes = r.bytes(r.word());
        s.buffers.push_back(std::move(b));
    }
    require(r.cursor == r.end, "snapshot has trailing bytes");
    s.validate();
    return s;
}

CountedState CountedState::select_colony(uint32_t colony) const {
    validate();
    require(colony < colonies, "snapshot colony out of range");
    CountedState out = *this;
    out.colonies = 1;
    out.tasks = {tasks[colony]};
    out.incumbents = {incumbents[colony]};
    for (size_t i = 0; i < buffers.size(); ++i) {
        const StateBuffer& source = buffers[i];
        StateBuffer& target = out.buffers[i];
        target.bytes.clear();
        const size_t stride = source.bytes.size() / source.planes / colonies;
        for (uint64_t plane = 0; plane < source.planes; ++plane) {
            const auto first = source.bytes.begin() + static_cast<std::ptrdiff_t>((plane * colonies + colony) * stride);
            target.bytes.insert(target.bytes.end(), first, first + static_cast<std::ptrdiff_t>(stride));
        }
    }
    out.validate();
    return out;
}

const StateBuffer& CountedState::buffer(const char* name) const {
    for (const auto& b : buffers)
        if (b.name == name) return b;
    throw std::invalid_argument("snapshot buffer missing");
}

}  // namespace faco
