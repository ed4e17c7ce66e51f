// CountedState container codec. The byte format (reference
// cpp/include/gp_faco/counted_state.hpp; its magic, word width, field order
// and structure rules are the interface snapshot files share) is:
//
//   word  magic            "GFACOST2" (0x3254534f43414647); an older
//                          "GFACOST1" container carries 8 more bytes at the
//                          very end, which a reader skips
//   word  dimension, colonies
//   word  ants, primary_width, backup_width, ls_width
//   real  beta, retention, p_best, epoch_source_probability
//   word  ls_evaluation_limit, initial_ls_evaluation_limit
//   word  experiment_mask, completed_batches, progress_evaluation_limit
//   list  tasks:       word instance_key, word seed
//   list  incumbents:  real cost, list of words (the tour)
//   list  buffers:     run name, word element_bytes, word planes, run bytes
//
// where `word` is a little-endian 64-bit unsigned integer, `real` the bit
// pattern of an IEEE double as a word, `list` a word count followed by the
// items and `run` a word byte count followed by the bytes.
//
// The format is walked by one function, transfer(), for every direction: the
// encoder reads each field and emits it, the size counter only adds up, and
// the decoder fills the field from the input after checking that the input
// still holds it (so no allocation exceeds what the input can back).
#include "counted_state.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace faco {
namespace {

constexpr uint64_t kMagic = 0x3254534f43414647ULL;        // "GFACOST2"
constexpr uint64_t kLegacyMagic = 0x3154534f43414647ULL;  // "GFACOST1"
constexpr size_t kLegacyTrailer = 8;
constexpr size_t kWord = 8;

// Limits of the structure rules (also bound what a decoder allocates).
constexpr uint64_t kMaxColonies = 128, kMaxAnts = 4096, kMaxBuffers = 64, kMaxName = 128;
constexpr uint64_t kMaxPlanes = 12, kMaxElement = 65536;

[[noreturn]] void reject(const char* why) { throw std::invalid_argument(why); }
void expect(bool ok, const char* why) { if (!ok) reject(why); }

uint64_t bits_of(double v) { uint64_t b; std::memcpy(&b, &v, sizeof b); return b; }
double real_of(uint64_t b) { double v; std::memcpy(&v, &b, sizeof v); return v; }

// Encoder: with a target, appends the bytes; without, only counts them.
class Emit {
public:
    explicit Emit(std::vector<uint8_t>* target) : out_(target) {}
    size_t size() const { return size_; }

    void word(uint64_t v) {
        if (out_) {
            uint8_t b[kWord];
            for (size_t i = 0; i < kWord; ++i) b[i] = static_cast<uint8_t>(v >> (8 * i));
            out_->insert(out_->end(), b, b + kWord);
        }
        size_ += kWord;
    }
    void bytes(const uint8_t* p, size_t n) {
        if (out_) out_->insert(out_->end(), p, p + n);
        size_ += n;
    }
    // the transfer() interface
    void narrow(uint32_t& v) { word(v); }
    void wide(uint64_t& v) { word(v); }
    void real(double& v) { word(bits_of(v)); }
    template<class T, class Item> void list(std::vector<T>& items, uint64_t, size_t, Item&& item) {
        word(items.size());
        for (T& x : items) item(x);
    }
    void text(std::string& s, uint64_t) {
        word(s.size());
        bytes(reinterpret_cast<const uint8_t*>(s.data()), s.size());
    }
    void blob(std::vector<uint8_t>& b) {
        word(b.size());
        bytes(b.data(), b.size());
    }

private:
    std::vector<uint8_t>* out_;
    size_t size_ = 0;
};

// Decoder over [next, last): every read first checks what is left.
class Parse {
public:
    Parse(const uint8_t* first, const uint8_t* last) : next_(first), last_(last) {}
    size_t left() const { return static_cast<size_t>(last_ - next_); }
    void drop_tail(size_t n) {
        expect(left() >= n, "snapshot file truncated");
        last_ -= n;
    }

    uint64_t word() {
        expect(left() >= kWord, "snapshot metadata truncated");
        uint64_t v = 0;
        for (size_t i = 0; i < kWord; ++i) v |= static_cast<uint64_t>(next_[i]) << (8 * i);
        next_ += kWord;
        return v;
    }
    // the transfer() interface
    void narrow(uint32_t& v) {
        const uint64_t w = word();
        expect(w <= UINT32_MAX, "snapshot integer overflow");
        v = static_cast<uint32_t>(w);
    }
    void wide(uint64_t& v) { v = word(); }
    void real(double& v) { v = real_of(word()); }
    // `limit` bounds the count; `least` is the fewest bytes one item takes,
    // so the count cannot promise more items than the input holds
    template<class T, class Item> void list(std::vector<T>& items, uint64_t limit, size_t least, Item&& item) {
        const uint64_t count = word();
        expect(count <= limit && count <= left() / least, "snapshot list count out of range");
        items.assign(static_cast<size_t>(count), T{});
        for (T& x : items) item(x);
    }
    void text(std::string& s, uint64_t limit) {
        const uint64_t length = word();
        expect(length <= limit, "snapshot buffer name out of range");
        expect(length <= left(), "snapshot buffer truncated");
        s.assign(reinterpret_cast<const char*>(next_), static_cast<size_t>(length));
        next_ += length;
    }
    void blob(std::vector<uint8_t>& b) {
        const uint64_t length = word();
        expect(length <= left(), "snapshot buffer truncated");
        b.assign(next_, next_ + length);
        next_ += length;
    }

private:
    const uint8_t* next_;
    const uint8_t* last_;
};

// Every field after the magic, in container order.
template<class Codec> void transfer(CountedState& s, Codec& io) {
    io.narrow(s.dimension);
    io.narrow(s.colonies);
    faco_settings& c = s.settings;
    io.narrow(c.ants);
    io.narrow(c.primary_width);
    io.narrow(c.backup_width);
    io.narrow(c.ls_width);
    io.real(c.beta);
    io.real(c.retention);
    io.real(c.p_best);
    io.real(c.epoch_source_probability);
    io.wide(c.ls_evaluation_limit);
    io.wide(c.initial_ls_evaluation_limit);
    io.narrow(s.experiment_mask);
    io.wide(s.completed_batches);
    io.wide(s.progress_evaluation_limit);
    io.list(s.tasks, kMaxColonies, 2 * kWord, [&](std::pair<uint64_t, uint64_t>& task) {
        io.wide(task.first);
        io.wide(task.second);
    });
    io.list(s.incumbents, kMaxColonies, 2 * kWord, [&](StateIncumbent& item) {
        io.real(item.cost);
        io.list(item.tour, UINT32_MAX, kWord, [&](uint32_t& city) { io.narrow(city); });
    });
    io.list(s.buffers, kMaxBuffers, 4 * kWord, [&](StateBuffer& b) {
        io.text(b.name, kMaxName);
        io.wide(b.element_bytes);
        io.wide(b.planes);
        io.blob(b.bytes);
    });
}

bool is_permutation_of_cities(const std::vector<uint32_t>& tour, uint32_t dimension) {
    if (tour.size() != dimension) return false;
    std::vector<uint8_t> seen(dimension, 0);
    for (uint32_t city : tour) {
        if (city >= dimension || seen[city]) return false;
        seen[city] = 1;
    }
    return true;
}

}  // namespace

bool same_settings(const faco_settings& a, const faco_settings& b) {
    // field by field: the struct has no padding bytes to compare, and the
    // reals compare as values (the snapshot keeps them bit for bit)
    return a.ants == b.ants && a.primary_width == b.primary_width && a.backup_width == b.backup_width &&
           a.ls_width == b.ls_width && a.beta == b.beta && a.retention == b.retention && a.p_best == b.p_best &&
           a.epoch_source_probability == b.epoch_source_probability &&
           a.ls_evaluation_limit == b.ls_evaluation_limit &&
           a.initial_ls_evaluation_limit == b.initial_ls_evaluation_limit;
}

// Structure rules of the container (the reference's, counted_state.cpp:57-76),
// checked section by section.
void CountedState::validate() const {
    // shape and tasks
    expect(dimension >= 3 && colonies >= 1 && colonies <= kMaxColonies, "snapshot shape or tasks invalid");
    expect(settings.ants >= 1 && settings.ants <= kMaxAnts, "snapshot shape or tasks invalid");
    expect(tasks.size() == colonies && incumbents.size() == colonies, "snapshot shape or tasks invalid");
    // host incumbents: a positive finite cost and a permutation of the cities
    for (const StateIncumbent& item : incumbents) {
        expect(item.tour.size() == dimension && std::isfinite(item.cost) && item.cost > 0,
               "snapshot host incumbent invalid");
        expect(is_permutation_of_cities(item.tour, dimension), "snapshot host tour is not a permutation");
    }
    // progress: whole ant batches, at most UINT32_MAX of them, not passed,
    // and a mask that keeps a hold action
    const uint64_t ants = settings.ants;
    const bool progress_ok = progress_evaluation_limit > 0 && progress_evaluation_limit % ants == 0 &&
                             progress_evaluation_limit / ants <= UINT32_MAX &&
                             completed_batches <= progress_evaluation_limit / ants;
    expect(progress_ok && (experiment_mask & 0xffffu) != 0, "snapshot batches, progress budget or mask invalid");
    // buffers: distinct non-empty names, element width and planes in range,
    // bytes a whole number of per-colony slices of every plane
    expect(!buffers.empty() && buffers.size() <= kMaxBuffers, "snapshot buffer set invalid");
    for (size_t i = 0; i < buffers.size(); ++i) {
        const StateBuffer& b = buffers[i];
        bool ok = !b.name.empty() && b.name.size() <= kMaxName && b.element_bytes >= 1 &&
                  b.element_bytes <= kMaxElement && b.planes >= 1 && b.planes <= kMaxPlanes && !b.bytes.empty() &&
                  b.bytes.size() % (b.element_bytes * b.planes * colonies) == 0;
        for (size_t j = 0; ok && j < i; ++j) ok = buffers[j].name != b.name;
        expect(ok, "snapshot buffer name, ABI or plane layout invalid");
    }
}

std::vector<uint8_t> CountedState::serialize() const {
    std::vector<uint8_t> out;
    out.reserve(serialized_size());
    Emit emit(&out);
    emit.word(kMagic);
    transfer(const_cast<CountedState&>(*this), emit);   // the encoder only reads
    return out;
}

size_t CountedState::serialized_size() const {
    Emit count(nullptr);
    count.word(kMagic);
    transfer(const_cast<CountedState&>(*this), count);
    return count.size();
}

CountedState CountedState::deserialize(const uint8_t* data, size_t size) {
    expect(data != nullptr && size >= 2 * kWord, "snapshot file truncated");
    Parse in(data, data + size);
    const uint64_t magic = in.word();
    expect(magic == kMagic || magic == kLegacyMagic, "unknown snapshot format");
    if (magic == kLegacyMagic) in.drop_tail(kLegacyTrailer);
    CountedState s;
    transfer(s, in);
    expect(in.left() == 0, "snapshot has trailing bytes");
    s.validate();
    return s;
}

// One colony's snapshot: each buffer holds `planes` planes of `colonies`
// equal slices; the cut keeps slice `colony` of every plane.
CountedState CountedState::select_colony(uint32_t colony) const {
    validate();
    expect(colony < colonies, "snapshot colony out of range");
    CountedState one;
    one.dimension = dimension;
    one.colonies = 1;
    one.settings = settings;
    one.experiment_mask = experiment_mask;
    one.completed_batches = completed_batches;
    one.progress_evaluation_limit = progress_evaluation_limit;
    one.tasks.push_back(tasks[colony]);
    one.incumbents.push_back(incumbents[colony]);
    for (const StateBuffer& b : buffers) {
        StateBuffer cut{b.name, b.element_bytes, b.planes, {}};
        const size_t slice = b.bytes.size() / (b.planes * colonies);
        cut.bytes.reserve(slice * b.planes);
        for (uint64_t plane = 0; plane < b.planes; ++plane) {
            const uint8_t* first = b.bytes.data() + (plane * colonies + colony) * slice;
            cut.bytes.insert(cut.bytes.end(), first, first + slice);
        }
        one.buffers.push_back(std::move(cut));
    }
    one.validate();
    return one;
}

const StateBuffer& CountedState::buffer(const char* name) const {
    for (const StateBuffer& b : buffers)
        if (b.name == name) return b;
    reject("snapshot buffer missing");
}

}  // namespace faco
