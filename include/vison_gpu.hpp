// vison_gpu.hpp -- header-only C++ façade over libvdf_hip.so (include/vdf.h).
//
// Mirrors the sort / group-by surface of the reference's vison::Dataframe
// (include/vison/vison.hpp:187-616) for device-resident columns:
//   sort_by<ASC, Soft>(n)                 transform/sort_by/sort_by.hpp:9-19,
//                                         sort_by_mt.hpp:10-180
//   transform_group_by<Function>(x, n_col, colname)
//                                         transform/transform_group_by/
//                                         transform_group_by.hpp:11-14
//   group_by(key_col, value_cols)         per-group results (keys ascending)
//   materialize()                         materialize/materialize_mt.hpp:10
//   read_csv(file_name, delim, header_name, str_context)
//                                         read/read_csv/read_csv.hpp:9-216
// Columns live in HBM; every operation runs in the HIP library.  Errors from
// the library are thrown as vison_gpu::Error (the reference prints to
// std::cerr and returns; a thrown error keeps the frame unchanged instead).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "vdf.h"

namespace vison_gpu {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(vdf_status s)
{
    if (s != VDF_OK) throw Error(s, vdf_last_error());
}
inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(VDF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// GroupFunction of types/supported_group_methods.inl:1-7; Gather + a min /
// max functor (create_value_col.inl:48-58) becomes Min / Max.
enum class GroupFunction { Occurence, Sum, Mean, Min, Max };

struct Char16 {
    char c[16];
};

// owning device buffer
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) : bytes_(bytes)
    {
        if (bytes) hip_check(hipMalloc(&p_, bytes), "hipMalloc");
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), bytes_(o.bytes_)
    {
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_;
            bytes_ = o.bytes_;
            o.p_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void* get() const { return p_; }
    size_t bytes() const { return bytes_; }

  private:
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
    }
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

// one library handle (scratch arena) per frame
class Handle {
  public:
    explicit Handle(int device = 0, uint64_t budget = 0) { check(vdf_create(&h_, device, budget)); }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    Handle(Handle&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    Handle& operator=(Handle&& o) noexcept
    {
        if (this != &o) {
            if (h_) (void)vdf_destroy(h_);
            h_ = o.h_;
            o.h_ = nullptr;
        }
        return *this;
    }
    ~Handle()
    {
        if (h_) (void)vdf_destroy(h_);
    }
    vdf_handle* get() const { return h_; }

  private:
    vdf_handle* h_ = nullptr;
};

template <typename T>
struct TypeCode;
template <> struct TypeCode<uint32_t> { static constexpr int key = VDF_KEY_U32, val = VDF_VAL_U32; static constexpr char letter = 'u'; };
template <> struct TypeCode<int32_t> { static constexpr int key = VDF_KEY_I32, val = VDF_VAL_I32; static constexpr char letter = 'i'; };
template <> struct TypeCode<uint64_t> { static constexpr int key = VDF_KEY_U64, val = VDF_VAL_U64; static constexpr char letter = 'u'; };
template <> struct TypeCode<int64_t> { static constexpr int key = VDF_KEY_I64, val = VDF_VAL_I64; static constexpr char letter = 'i'; };
template <> struct TypeCode<double> { static constexpr int key = VDF_KEY_F64, val = VDF_VAL_F64; static constexpr char letter = 'd'; };
template <> struct TypeCode<float> { static constexpr int key = VDF_KEY_F32, val = -1; static constexpr char letter = 'd'; };
template <> struct TypeCode<Char16> { static constexpr int key = VDF_KEY_CHAR16, val = -1; static constexpr char letter = 'c'; };

// A device column: element size, key / value type codes, name.  A string
// column holds its strings padded with 0x00 to the longest (elem bytes each),
// as the reference's flat buffer for sorting strings (sort_string.inl:30-85).
struct Column {
    std::string name;
    size_t elem = 0;
    int key_type = -1;
    int val_type = -1;
    char letter = '?';
    bool str = false;
    DeviceBuffer data;
};

template <typename T>
struct GroupResult {
    size_t n_groups = 0;
    std::vector<T> keys;                 // ascending
    std::vector<uint32_t> offsets;       // n_groups + 1, into row_view
    std::vector<uint32_t> counts;
    std::vector<uint32_t> row_view;      // group-contiguous row ids
    // Indexed by value-column position: for a float64 column c, sum_f[c],
    // min_f[c], max_f[c] hold its groups (sum_i[c] ... empty); for an integer
    // column the *_i vectors do.  mean[c] is filled for every column.
    std::vector<std::vector<double>> sum_f, mean, min_f, max_f;
    std::vector<std::vector<int64_t>> sum_i, min_i, max_i;
};

class Dataframe {
  public:
    explicit Dataframe(int device = 0) : h_(device) {}

    unsigned int get_nrow() const { return nrow_; }
    unsigned int get_ncol() const { return static_cast<unsigned int>(cols_.size()); }
    bool in_view() const { return in_view_; }
    const std::vector<size_t>& get_row_view_idx() const { return row_view_idx_; }

    template <typename T>
    void add_col(const std::vector<T>& host, const std::string& name)
    {
        if (!cols_.empty() && host.size() != nrow_) throw Error(VDF_ERR_INVALID, "add_col: row count mismatch");
        Column c;
        c.name = name;
        c.elem = sizeof(T);
        c.key_type = TypeCode<T>::key;
        c.val_type = TypeCode<T>::val;
        c.letter = TypeCode<T>::letter;
        c.data = DeviceBuffer(host.size() * sizeof(T));
        if (!host.empty())
            hip_check(hipMemcpy(c.data.get(), host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice),
                      "hipMemcpy H2D");
        nrow_ = static_cast<unsigned int>(host.size());
        cols_.push_back(std::move(c));
    }

    // std::string column: stored padded (PaddingChar 0x00); sort_by orders it
    // bytewise as signed chars, as sort_string does.
    void add_col(const std::vector<std::string>& host, const std::string& name)
    {
        if (!cols_.empty() && host.size() != nrow_) throw Error(VDF_ERR_INVALID, "add_col: row count mismatch");
        size_t w = 1;
        for (const std::string& x : host) w = std::max(w, x.size());
        if (w > VDF_MAX_STRING_BYTES) throw Error(VDF_ERR_INVALID, "add_col: string longer than VDF_MAX_STRING_BYTES");
        std::vector<char> flat(host.size() * w, 0);
        for (size_t i = 0; i < host.size(); ++i) std::memcpy(flat.data() + i * w, host[i].data(), host[i].size());
        Column c;
        c.name = name;
        c.elem = w;
        c.letter = 's';
        c.str = true;
        c.data = DeviceBuffer(flat.size());
        if (!flat.empty())
            hip_check(hipMemcpy(c.data.get(), flat.data(), flat.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
        nrow_ = static_cast<unsigned int>(host.size());
        cols_.push_back(std::move(c));
    }

    // the strings of a string column (trailing padding removed)
    std::vector<std::string> get_str_col(unsigned int n) const
    {
        const Column& c = col(n);
        if (!c.str) throw Error(VDF_ERR_INVALID, "get_str_col: not a string column");
        std::vector<char> flat(static_cast<size_t>(nrow_) * c.elem);
        if (!flat.empty())
            hip_check(hipMemcpy(flat.data(), c.data.get(), flat.size(), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        std::vector<std::string> out(nrow_);
        for (size_t i = 0; i < nrow_; ++i) {
            size_t len = c.elem;
            while (len > 0 && flat[i * c.elem + len - 1] == 0) --len;
            out[i].assign(flat.data() + i * c.elem, len);
        }
        return out;
    }

    template <typename T>
    std::vector<T> get_col(unsigned int n) const
    {
        const Column& c = col(n);
        if (c.elem != sizeof(T)) throw Error(VDF_ERR_INVALID, "get_col: element size mismatch");
        std::vector<T> out(nrow_);
        if (nrow_)
            hip_check(hipMemcpy(out.data(), c.data.get(), nrow_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return out;
    }

    // Reorder rows by column n (sort_by.hpp:9-19).  ASC=false is the
    // reference's descending path: the ascending order reversed, so equal keys
    // come out in reverse row order (sort_uintegers.inl:216-218).  Soft=true keeps the
    // permutation as a row view (sort_by_mt.hpp:160-177), Soft=false gathers
    // every column (permute_block_mt.inl:39-41).
    template <bool ASC = true, bool Soft = true>
    void sort_by(unsigned int& n)
    {
        if (!Soft && in_view_) throw Error(VDF_ERR_INVALID, "sort_by: materialize() the view first");
        const Column& k = col(n);
        if (k.key_type < 0 && !k.str) throw Error(VDF_ERR_UNSUPPORTED, "sort_by: column type has no radix key path");
        if (nrow_ == 0) return;
        DeviceBuffer perm(static_cast<size_t>(nrow_) * 4);
        if constexpr (!Soft) {
            // hard sort of a numeric key: the first other 4- or 8-byte column
            // rides through the radix passes (vdf_sort_by) and the sorted keys
            // come out of the last pass; only the remaining columns are gathered
            int carry = -1;
            for (size_t c = 0; c < cols_.size() && !k.str; ++c)
                if (c != n && !cols_[c].str && (cols_[c].elem == 4 || cols_[c].elem == 8)) {
                    carry = static_cast<int>(c);
                    break;
                }
            if (carry >= 0) {
                Column& cv = cols_[static_cast<size_t>(carry)];
                DeviceBuffer skeys(k.data.bytes()), svals(cv.data.bytes());
                check(vdf_sort_by(h_.get(), k.key_type, k.data.get(), nrow_, ASC ? 0u : VDF_SORT_DESC_REVERSE,
                                  perm.get(), skeys.get(), static_cast<uint32_t>(cv.elem), cv.data.get(), svals.get(),
                                  nullptr));
                permute_all(perm.get(), {n, static_cast<unsigned int>(carry)});
                cols_[n].data = std::move(skeys);
                cv.data = std::move(svals);
                return;
            }
        }
        // In view mode the key is sorted as the view presents it (gathered
        // through the view), so the new view orders the viewed rows by the key,
        // stably: new_view[j] = view[q[j]] for the sort q of the viewed keys.
        // The reference instead scatters the stored-order permutation p into
        // the view by cycle swaps (sort_by_mt.hpp:164-173), new_view[p[i]] =
        // view[i], which does not order the rows by the key.
        const void* keys = k.data.get();
        DeviceBuffer viewed;
        if (Soft && in_view_) {
            std::vector<uint32_t> v(row_view_idx_.begin(), row_view_idx_.end());
            DeviceBuffer vp(v.size() * 4);
            hip_check(hipMemcpy(vp.get(), v.data(), v.size() * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
            viewed = DeviceBuffer(k.data.bytes());
            check(vdf_gather(h_.get(), static_cast<uint32_t>(k.elem), k.data.get(), vp.get(), 0, nrow_, viewed.get(),
                             nullptr));
            keys = viewed.get();
        }
        if (k.str)  // sort_string (sort_by_mt.hpp:72-115 dispatch)
            check(vdf_argsort_bytes(h_.get(), keys, static_cast<uint32_t>(k.elem), nrow_,
                                    ASC ? 0u : VDF_SORT_DESC_REVERSE, perm.get(), nullptr, nullptr));
        else
            check(vdf_argsort(h_.get(), k.key_type, keys, nrow_, ASC ? 0u : VDF_SORT_DESC_REVERSE, perm.get(), nullptr,
                              nullptr));
        if constexpr (Soft) {
            std::vector<uint32_t> p(nrow_);
            hip_check(hipMemcpy(p.data(), perm.get(), nrow_ * 4ull, hipMemcpyDeviceToHost), "hipMemcpy D2H");
            if (in_view_) {  // compose with the existing view (gather form)
                std::vector<size_t> r(nrow_);
                for (size_t i = 0; i < nrow_; ++i) r[i] = row_view_idx_[p[i]];
                row_view_idx_.swap(r);
            } else {
                row_view_idx_.assign(p.begin(), p.end());
            }
            in_view_ = true;
        } else {
            permute_all(perm.get());
        }
    }

    // Apply the row view to every column (materialize_mt.hpp:10).
    void materialize()
    {
        if (!in_view_) return;
        std::vector<uint32_t> p(row_view_idx_.begin(), row_view_idx_.end());
        DeviceBuffer perm(p.size() * 4);
        if (!p.empty())
            hip_check(hipMemcpy(perm.get(), p.data(), p.size() * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        permute_all(perm.get());
        row_view_idx_.clear();
        in_view_ = false;
    }

    // Device-resident results of one group-by (per-group arrays sized n).
    struct DeviceGroups {
        size_t n = 0;
        DeviceBuffer ng, off, cnt, perm;
        std::vector<DeviceBuffer> keys;  // one per key column
        std::vector<DeviceBuffer> agg;   // 4 per value column: sum, mean, min, max
        vdf_groupby_out out{};
        uint64_t n_groups() const
        {
            uint64_t g = 0;
            hip_check(hipMemcpy(&g, ng.get(), 8, hipMemcpyDeviceToHost), "hipMemcpy D2H");
            return g;
        }
    };

    // Group by one or more key columns (lexicographic, first column most
    // significant) and aggregate value columns; results stay on the device.
    DeviceGroups group_by_device(const std::vector<unsigned int>& key_cols, const std::vector<unsigned int>& vals,
                                 uint32_t aggs = VDF_AGG_ALL)
    {
        if (key_cols.empty() || key_cols.size() > VDF_MAX_KEY_COLS)
            throw Error(VDF_ERR_INVALID, "group_by: 1 .. VDF_MAX_KEY_COLS key columns");
        if (vals.size() > VDF_MAX_VALUE_COLS) throw Error(VDF_ERR_INVALID, "group_by: too many value columns");
        const size_t n = nrow_;
        DeviceGroups D;
        D.n = n;
        const size_t cap = std::max<size_t>(n, 1);
        D.ng = DeviceBuffer(8);
        D.off = DeviceBuffer((n + 1) * 4);
        D.cnt = DeviceBuffer(cap * 4);
        D.perm = DeviceBuffer(cap * 4);
        std::vector<vdf_key_col> kc;
        std::vector<void*> kout;
        for (unsigned int k : key_cols) {
            const Column& c = col(k);
            if (c.key_type < 0 && !c.str) throw Error(VDF_ERR_UNSUPPORTED, "group_by: column type has no key path");
            // a std::string column (stored padded) is a byte-string key column
            kc.push_back(c.str ? vdf_key_col{VDF_KEY_BYTES, c.data.get(), static_cast<uint32_t>(c.elem)}
                               : vdf_key_col{c.key_type, c.data.get(), 0});
            D.keys.emplace_back(cap * c.elem);
            kout.push_back(D.keys.back().get());
        }
        vdf_groupby_out& out = D.out;
        out.n_groups = static_cast<uint64_t*>(D.ng.get());
        out.group_keys = kout[0];
        out.offsets = static_cast<uint32_t*>(D.off.get());
        out.counts = static_cast<uint32_t*>(D.cnt.get());
        out.perm = static_cast<uint32_t*>(D.perm.get());
        std::vector<vdf_value_col> vc;
        for (size_t c = 0; c < vals.size(); ++c) {
            const Column& v = col(vals[c]);
            if (v.val_type < 0) throw Error(VDF_ERR_UNSUPPORTED, "group_by: value column type");
            vc.push_back(vdf_value_col{v.val_type, v.data.get()});
            for (int a = 0; a < 4; ++a) D.agg.emplace_back(cap * 8);
            out.aggs[c].sum = D.agg[4 * c + 0].get();
            out.aggs[c].mean = static_cast<double*>(D.agg[4 * c + 1].get());
            out.aggs[c].min = D.agg[4 * c + 2].get();
            out.aggs[c].max = D.agg[4 * c + 3].get();
        }
        const Column& k0 = col(key_cols[0]);
        if (k0.str && kc.size() == 1)  // a std::string key column
            check(vdf_groupby_bytes(h_.get(), k0.data.get(), static_cast<uint32_t>(k0.elem), n, vc.data(),
                                    static_cast<int>(vc.size()), aggs, &out, nullptr));
        else if (kc.size() == 1)
            check(vdf_groupby(h_.get(), kc[0].type, kc[0].data, n, vc.data(), static_cast<int>(vc.size()), aggs, &out,
                              nullptr));
        else
            check(vdf_groupby_multi(h_.get(), kc.data(), static_cast<int>(kc.size()), n, vc.data(),
                                    static_cast<int>(vc.size()), aggs, &out, kout.data(), nullptr));
        return D;
    }

    // Group by key column `key` and aggregate value columns (ascending keys).
    template <typename K>
    GroupResult<K> group_by(unsigned int key, const std::vector<unsigned int>& vals, uint32_t aggs = VDF_AGG_ALL)
    {
        const Column& k = col(key);
        if (k.key_type < 0 || k.elem != sizeof(K)) throw Error(VDF_ERR_INVALID, "group_by: bad key column");
        DeviceGroups D = group_by_device({key}, vals, aggs);
        return to_host<K>(D, vals);
    }

    // Group by a std::string key column: groups in ascending order of the
    // strings' bytes as signed chars (keys without their padding).
    GroupResult<std::string> group_by_str(unsigned int key, const std::vector<unsigned int>& vals,
                                          uint32_t aggs = VDF_AGG_ALL)
    {
        const Column& k = col(key);
        if (!k.str) throw Error(VDF_ERR_INVALID, "group_by_str: not a string column");
        DeviceGroups D = group_by_device({key}, vals, aggs);
        GroupResult<uint8_t> r = to_host<uint8_t>(D, vals, /*with_keys=*/false);
        GroupResult<std::string> R;
        R.n_groups = r.n_groups;
        std::vector<char> flat(r.n_groups * k.elem);
        if (!flat.empty())
            hip_check(hipMemcpy(flat.data(), D.keys[0].get(), flat.size(), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        R.keys.resize(r.n_groups);
        for (size_t g = 0; g < r.n_groups; ++g) {
            size_t len = k.elem;
            while (len > 0 && flat[g * k.elem + len - 1] == 0) --len;
            R.keys[g].assign(flat.data() + g * k.elem, len);
        }
        R.offsets = std::move(r.offsets);
        R.counts = std::move(r.counts);
        R.row_view = std::move(r.row_view);
        R.sum_f = std::move(r.sum_f);
        R.mean = std::move(r.mean);
        R.min_f = std::move(r.min_f);
        R.max_f = std::move(r.max_f);
        R.sum_i = std::move(r.sum_i);
        R.min_i = std::move(r.min_i);
        R.max_i = std::move(r.max_i);
        return R;
    }

    // Group by several key columns; group keys of column i in keys_of[i] (raw bytes).
    struct MultiGroupResult {
        GroupResult<uint8_t> r;                    // r.keys unused
        std::vector<std::vector<uint8_t>> keys_of;  // per key column, n_groups cells
    };
    MultiGroupResult group_by_multi(const std::vector<unsigned int>& key_cols, const std::vector<unsigned int>& vals,
                                    uint32_t aggs = VDF_AGG_ALL)
    {
        DeviceGroups D = group_by_device(key_cols, vals, aggs);
        MultiGroupResult M;
        M.r = to_host<uint8_t>(D, vals, /*with_keys=*/false);
        for (size_t c = 0; c < key_cols.size(); ++c) {
            const size_t eb = col(key_cols[c]).elem;
            M.keys_of.emplace_back(M.r.n_groups * eb);
            if (M.r.n_groups)
                hip_check(hipMemcpy(M.keys_of.back().data(), D.keys[c].get(), M.r.n_groups * eb,
                                    hipMemcpyDeviceToHost),
                          "hipMemcpy D2H");
        }
        return M;
    }

    // transform_group_by (transform_group_by.hpp:11): append a per-row float64
    // column holding the aggregate of the row's group, as CreateValueCol does
    // (create_value_col.inl:36-61), for one or more key columns x.  The
    // group-by and the broadcast both run on the device.  Mean = sum / count.
    template <GroupFunction F>
    void transform_group_by(const std::vector<unsigned int>& x, unsigned int n_col, const std::string& colname = "n")
    {
        const std::vector<unsigned int> vals = (F == GroupFunction::Occurence) ? std::vector<unsigned int>{}
                                                                               : std::vector<unsigned int>{n_col};
        DeviceGroups D = group_by_device(x, vals);
        const void* src = nullptr;
        int src_type = VDF_VAL_U32;
        if constexpr (F == GroupFunction::Occurence) {
            src = D.cnt.get();
        } else {
            const Column& v = col(n_col);
            const bool f = v.val_type == VDF_VAL_F64;
            const bool u = v.val_type == VDF_VAL_U32 || v.val_type == VDF_VAL_U64;
            if (F == GroupFunction::Sum) {
                src = D.agg[0].get();
                src_type = f ? VDF_VAL_F64 : (u ? VDF_VAL_U64 : VDF_VAL_I64);
            } else if (F == GroupFunction::Mean) {
                src = D.agg[1].get();
                src_type = VDF_VAL_F64;
            } else {
                src = D.agg[F == GroupFunction::Min ? 2 : 3].get();
                src_type = v.val_type;
            }
        }
        Column c;
        c.name = colname;
        c.elem = 8;
        c.key_type = VDF_KEY_F64;
        c.val_type = VDF_VAL_F64;
        c.letter = 'd';
        c.data = DeviceBuffer(static_cast<size_t>(nrow_) * 8);
        check(vdf_broadcast(h_.get(), src_type, VDF_VAL_F64, src, static_cast<const uint32_t*>(D.off.get()),
                            static_cast<const uint64_t*>(D.ng.get()), std::max<size_t>(nrow_, 1),
                            static_cast<const uint32_t*>(D.perm.get()), nrow_, c.data.get(), nullptr));
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        cols_.push_back(std::move(c));
    }

    // Load a CSV file into this (empty) frame: the file's bytes go to HBM once
    // and vdf_read_csv parses them there (read_csv.hpp:9-216; types as its
    // classification of the first 10 rows).  IntT / UIntT columns are
    // int64_t / uint64_t (int_bytes 8), CharT cells Char16 (char_width 16,
    // padded with TrailingChar), bool columns uint32_t 0 / 1, other text
    // columns std::string columns.  The template parameters are the
    // reference's (read_csv.hpp:3-8): the frame keeps data rows strt_row ..
    // end_row - 1 (end_row 0: to the last row), as standard_parser.inl:18-58
    // slices them; CORES, WARMING and MEM_CLEAN choose among its CPU parsers
    // and have no effect here.
    template <unsigned int strt_row = 0, unsigned int end_row = 0, unsigned int CORES = 1, bool WARMING = false,
              bool MEM_CLEAN = false, char TrailingChar = '0'>
    void read_csv(const std::string& file_name, char delim = ',', bool header_name = true, char str_context = '\'')
    {
        static_assert(end_row == 0 || strt_row < end_row, "read_csv: strt_row must be lower than end_row");
        if (!cols_.empty()) throw Error(VDF_ERR_INVALID, "read_csv: the frame already has columns");
        std::ifstream in(file_name, std::ios::binary | std::ios::ate);
        if (!in) throw Error(VDF_ERR_INVALID, "read_csv: cannot open " + file_name);
        const std::streamsize size = in.tellg();
        in.seekg(0);
        std::vector<char> host(static_cast<size_t>(std::max<std::streamsize>(size, 0)));
        if (size > 0 && !in.read(host.data(), size)) throw Error(VDF_ERR_INVALID, "read_csv: cannot read " + file_name);
        DeviceBuffer text(host.size());
        if (!host.empty())
            hip_check(hipMemcpy(text.get(), host.data(), host.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
        vdf_csv_options opt{};
        opt.delim = delim;
        opt.quote = str_context;
        opt.trailing = TrailingChar;
        opt.header = header_name ? 1 : 0;
        opt.int_bytes = 8;
        opt.char_width = 16;
        vdf_csv_table* t = nullptr;
        check(vdf_read_csv(h_.get(), text.get(), host.size(), &opt, &t));
        struct Free {
            vdf_csv_table* t;
            ~Free() { (void)vdf_csv_free(t); }
        } guard{t};
        uint64_t all = 0;
        int ncol = 0;
        check(vdf_csv_info(t, &all, &ncol));
        const uint64_t first = std::min<uint64_t>(strt_row, all);
        const uint64_t last = end_row == 0 ? all : std::min<uint64_t>(end_row, all);
        const uint64_t nrow = last > first ? last - first : 0;
        for (int k = 0; k < ncol; ++k) {
            vdf_csv_column vc{};
            check(vdf_csv_column_get(t, k, &vc));
            Column c;
            c.name = vc.name ? vc.name : "";
            c.elem = vc.width;
            switch (vc.kind) {
            case VDF_CSV_DOUBLE: set_types<double>(c); break;
            case VDF_CSV_INT: set_types<int64_t>(c); break;
            case VDF_CSV_UINT: set_types<uint64_t>(c); break;
            case VDF_CSV_BOOL: set_types<uint32_t>(c); break;
            case VDF_CSV_CHAR: set_types<Char16>(c); break;
            default:
                c.letter = 's';
                c.str = true;
                break;
            }
            c.data = DeviceBuffer(static_cast<size_t>(nrow) * c.elem);
            if (nrow)
                hip_check(hipMemcpy(c.data.get(), static_cast<const char*>(vc.data) + first * c.elem,
                                    static_cast<size_t>(nrow) * c.elem, hipMemcpyDeviceToDevice),
                          "hipMemcpy D2D");
            cols_.push_back(std::move(c));
        }
        nrow_ = static_cast<unsigned int>(nrow);
    }

    const std::string& get_colname(unsigned int n) const { return col(n).name; }

  private:
    template <typename T>
    static void set_types(Column& c)
    {
        c.key_type = TypeCode<T>::key;
        c.val_type = TypeCode<T>::val;
        c.letter = TypeCode<T>::letter;
    }

    template <typename K>
    GroupResult<K> to_host(const DeviceGroups& D, const std::vector<unsigned int>& vals, bool with_keys = true) const
    {
        const size_t n = D.n;
        GroupResult<K> R;
        const uint64_t g = D.n_groups();
        R.n_groups = g;
        if (with_keys) R.keys.resize(g);
        R.offsets.resize(g + 1);
        R.counts.resize(g);
        R.row_view.resize(n);
        for (auto* v : {&R.sum_f, &R.mean, &R.min_f, &R.max_f}) v->resize(vals.size());
        for (auto* v : {&R.sum_i, &R.min_i, &R.max_i}) v->resize(vals.size());
        auto d2h = [](void* dst, const void* src, size_t bytes) {
            if (bytes) hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        };
        if (with_keys) d2h(R.keys.data(), D.keys[0].get(), g * sizeof(K));
        d2h(R.offsets.data(), D.off.get(), (g + 1) * 4);
        d2h(R.counts.data(), D.cnt.get(), g * 4);
        d2h(R.row_view.data(), D.perm.get(), n * 4);
        for (size_t c = 0; c < vals.size(); ++c) {
            const Column& v = col(vals[c]);
            const vdf_agg_out& a = D.out.aggs[c];
            if (v.val_type == VDF_VAL_F64) {
                std::vector<double> s(g), m(g), lo(g), hi(g);
                d2h(s.data(), a.sum, g * 8);
                d2h(m.data(), a.mean, g * 8);
                d2h(lo.data(), a.min, g * 8);
                d2h(hi.data(), a.max, g * 8);
                R.sum_f[c] = std::move(s);
                R.mean[c] = std::move(m);
                R.min_f[c] = std::move(lo);
                R.max_f[c] = std::move(hi);
            } else {
                std::vector<int64_t> s(g), lo(g), hi(g);
                std::vector<double> m(g);
                d2h(s.data(), a.sum, g * 8);
                d2h(m.data(), a.mean, g * 8);
                const bool w4 = v.elem == 4;
                std::vector<uint32_t> l4(w4 ? g : 0), h4(w4 ? g : 0);
                if (w4) {
                    d2h(l4.data(), a.min, g * 4);
                    d2h(h4.data(), a.max, g * 4);
                    for (size_t i = 0; i < g; ++i) {
                        lo[i] = v.val_type == VDF_VAL_I32 ? static_cast<int32_t>(l4[i]) : static_cast<int64_t>(l4[i]);
                        hi[i] = v.val_type == VDF_VAL_I32 ? static_cast<int32_t>(h4[i]) : static_cast<int64_t>(h4[i]);
                    }
                } else {
                    d2h(lo.data(), a.min, g * 8);
                    d2h(hi.data(), a.max, g * 8);
                }
                R.sum_i[c] = std::move(s);
                R.mean[c] = std::move(m);
                R.min_i[c] = std::move(lo);
                R.max_i[c] = std::move(hi);
            }
        }
        return R;
    }

    const Column& col(unsigned int n) const
    {
        if (n >= cols_.size()) throw Error(VDF_ERR_INVALID, "column index out of range");
        return cols_[n];
    }

    // gather every column through perm, except the columns in `skip`
    void permute_all(void* perm, std::initializer_list<unsigned int> skip = {})
    {
        for (size_t i = 0; i < cols_.size(); ++i) {
            if (std::find(skip.begin(), skip.end(), static_cast<unsigned int>(i)) != skip.end()) continue;
            Column& c = cols_[i];
            DeviceBuffer dst(c.data.bytes());
            check(vdf_gather(h_.get(), static_cast<uint32_t>(c.elem), c.data.get(), perm, 0, nrow_, dst.get(),
                             nullptr));
            c.data = std::move(dst);
        }
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }

    Handle h_;
    unsigned int nrow_ = 0;
    std::vector<Column> cols_;
    bool in_view_ = false;
    std::vector<size_t> row_view_idx_;
};

}  // namespace vison_gpu
