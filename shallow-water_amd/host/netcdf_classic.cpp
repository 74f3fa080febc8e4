// netcdf_classic.cpp -- see netcdf_classic.hpp.
//
// File layout (NetCDF classic format specification, version 2 = 64-bit offsets):
//   "CDF" 0x02, numrecs (uint32), dim_list, gatt_list, var_list, then the fixed-size
//   variables at their begin offsets, then the records (each record holds one slice of every
//   record variable, in definition order).  Lists are ABSENT (two zero words) or
//   tag + count + entries; names are count + bytes padded to 4; all numbers big-endian.
#include "netcdf_classic.hpp"

#include <cmath>
#include <cstring>
#include <filesystem>
#include <stdexcept>

namespace {

constexpr uint32_t kDimTag = 0x0A, kVarTag = 0x0B, kAttTag = 0x0C;

void put_u32(std::vector<unsigned char> &o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back((unsigned char)(v >> s));
}
void put_u64(std::vector<unsigned char> &o, uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) o.push_back((unsigned char)(v >> s));
}
void put_name(std::vector<unsigned char> &o, const std::string &n) {
    put_u32(o, (uint32_t)n.size());
    o.insert(o.end(), n.begin(), n.end());
    while (o.size() % 4) o.push_back(0);
}
uint64_t be64(double d) {
    uint64_t u;
    std::memcpy(&u, &d, 8);
    return __builtin_bswap64(u);
}
uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

}  // namespace

NcClassicFile::NcClassicFile(const std::string &path) : path_(path) {
    const std::filesystem::path p(path);
    if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path());
    f_ = std::fopen(path.c_str(), "w+b");
    if (!f_) throw std::runtime_error("cannot create NetCDF file '" + path + "'");
}

NcClassicFile::~NcClassicFile() {
    if (f_) {
        if (!header_written_) {
            try {
                end_define();
            } catch (...) {
            }
        }
        std::fclose(f_);
    }
}

int NcClassicFile::add_dim(const std::string &name, std::size_t size) {
    if (header_written_) throw std::runtime_error("NetCDF: dimension after end of definition");
    if (size == 0)
        for (const Dim &d : dims_)
            if (d.size == 0) throw std::runtime_error("NetCDF: only one record dimension");
    dims_.push_back({name, size});
    return (int)dims_.size() - 1;
}

int NcClassicFile::add_var(const std::string &name, Type type, const std::vector<int> &dims) {
    if (header_written_) throw std::runtime_error("NetCDF: variable after end of definition");
    Var v;
    v.name = name;
    v.type = type;
    v.dims = dims;
    std::size_t n = 1;
    for (std::size_t k = 0; k < dims.size(); ++k) {
        const Dim &d = dims_.at((std::size_t)dims[k]);
        if (d.size == 0) {
            if (k != 0) throw std::runtime_error("NetCDF: the record dimension must come first");
            v.record = true;
        } else {
            n *= d.size;
        }
    }
    v.nelem_slice = n;
    v.vsize = pad4(n * type_size(type));
    if (v.vsize >= (uint64_t(1) << 32)) throw std::runtime_error("NetCDF: variable too large");
    vars_.push_back(v);
    return (int)vars_.size() - 1;
}

void NcClassicFile::put_att(int var, const std::string &name, const std::string &text) {
    Att a{name, NC_CHAR, std::vector<unsigned char>(text.begin(), text.end()),
          (uint32_t)text.size()};
    (var < 0 ? gatts_ : vars_.at((std::size_t)var).atts).push_back(a);
}

void NcClassicFile::put_att(int var, const std::string &name, int32_t value) {
    Att a{name, NC_INT, {}, 1};
    put_u32(a.bytes, (uint32_t)value);
    (var < 0 ? gatts_ : vars_.at((std::size_t)var).atts).push_back(a);
}

void NcClassicFile::put_att(int var, const std::string &name, double value) {
    Att a{name, NC_DOUBLE, {}, 1};
    uint64_t u;
    std::memcpy(&u, &value, 8);
    put_u64(a.bytes, u);
    (var < 0 ? gatts_ : vars_.at((std::size_t)var).atts).push_back(a);
}

void NcClassicFile::end_define() {
    if (header_written_) return;
    auto att_list = [](std::vector<unsigned char> &o, const std::vector<Att> &atts) {
        if (atts.empty()) {
            put_u32(o, 0);
            put_u32(o, 0);
            return;
        }
        put_u32(o, kAttTag);
        put_u32(o, (uint32_t)atts.size());
        for (const Att &a : atts) {
            put_name(o, a.name);
            put_u32(o, (uint32_t)a.type);
            put_u32(o, a.nelems);
            o.insert(o.end(), a.bytes.begin(), a.bytes.end());
            while (o.size() % 4) o.push_back(0);
        }
    };
    // the header size does not depend on the begin offsets (fixed 8-byte fields), so build
    // it once with zeros to learn its length, then with the real offsets
    auto build = [&](std::vector<unsigned char> &o) {
        o.clear();
        const char magic[4] = {'C', 'D', 'F', 2};
        o.insert(o.end(), magic, magic + 4);
        put_u32(o, (uint32_t)numrecs_);
        if (dims_.empty()) {
            put_u32(o, 0);
            put_u32(o, 0);
        } else {
            put_u32(o, kDimTag);
            put_u32(o, (uint32_t)dims_.size());
            for (const Dim &d : dims_) {
                put_name(o, d.name);
                put_u32(o, (uint32_t)d.size);
            }
        }
        att_list(o, gatts_);
        if (vars_.empty()) {
            put_u32(o, 0);
            put_u32(o, 0);
        } else {
            put_u32(o, kVarTag);
            put_u32(o, (uint32_t)vars_.size());
            for (const Var &v : vars_) {
                put_name(o, v.name);
                put_u32(o, (uint32_t)v.dims.size());
                for (int d : v.dims) put_u32(o, (uint32_t)d);
                att_list(o, v.atts);
                put_u32(o, (uint32_t)v.type);
                put_u32(o, (uint32_t)v.vsize);
                put_u64(o, v.begin);
            }
        }
    };
    std::vector<unsigned char> hdr;
    build(hdr);
    uint64_t off = hdr.size();
    for (Var &v : vars_)
        if (!v.record) {
            v.begin = off;
            off += v.vsize;
        }
    recsize_ = 0;
    for (Var &v : vars_)
        if (v.record) {
            v.begin = off + recsize_;
            recsize_ += v.vsize;
        }
    build(hdr);
    write_at(0, hdr.data(), hdr.size());
    // fixed variables default to zeros until written
    std::vector<unsigned char> zeros(1 << 16, 0);
    for (const Var &v : vars_)
        if (!v.record)
            for (uint64_t done = 0; done < v.vsize;) {
                const std::size_t n = (std::size_t)std::min<uint64_t>(zeros.size(), v.vsize - done);
                write_at(v.begin + done, zeros.data(), n);
                done += n;
            }
    header_written_ = true;
    std::fflush(f_);
}

void NcClassicFile::write_at(uint64_t off, const void *p, std::size_t n) {
    if (fseeko(f_, (off_t)off, SEEK_SET) != 0 || std::fwrite(p, 1, n, f_) != n)
        throw std::runtime_error("NetCDF: write failed on '" + path_ + "'");
}

void NcClassicFile::update_numrecs(std::size_t rec) {
    if (rec + 1 > numrecs_) {
        numrecs_ = rec + 1;
        unsigned char b[4] = {(unsigned char)(numrecs_ >> 24), (unsigned char)(numrecs_ >> 16),
                              (unsigned char)(numrecs_ >> 8), (unsigned char)numrecs_};
        write_at(4, b, 4);
    }
}

void NcClassicFile::put_fixed(int var, const double *data) {
    const Var &v = vars_.at((std::size_t)var);
    if (!header_written_ || v.record || v.type != NC_DOUBLE)
        throw std::runtime_error("NetCDF: put_fixed on '" + v.name + "'");
    std::vector<uint64_t> buf(v.nelem_slice);
    for (std::size_t k = 0; k < buf.size(); ++k) buf[k] = be64(data[k]);
    write_at(v.begin, buf.data(), buf.size() * 8);
}

void NcClassicFile::put_record(int var, std::size_t rec, const double *data) {
    const Var &v = vars_.at((std::size_t)var);
    if (!header_written_ || !v.record || v.type != NC_DOUBLE)
        throw std::runtime_error("NetCDF: put_record on '" + v.name + "'");
    std::vector<uint64_t> buf(v.nelem_slice);
    for (std::size_t k = 0; k < buf.size(); ++k) buf[k] = be64(data[k]);
    write_at(v.begin + rec * recsize_, buf.data(), buf.size() * 8);
    update_numrecs(rec);
    std::fflush(f_);
}

void NcClassicFile::put_record(int var, std::size_t rec, const int32_t *data) {
    const Var &v = vars_.at((std::size_t)var);
    if (!header_written_ || !v.record || v.type != NC_INT)
        throw std::runtime_error("NetCDF: put_record on '" + v.name + "'");
    std::vector<uint32_t> buf(v.nelem_slice);
    for (std::size_t k = 0; k < buf.size(); ++k) buf[k] = __builtin_bswap32((uint32_t)data[k]);
    write_at(v.begin + rec * recsize_, buf.data(), buf.size() * 4);
    update_numrecs(rec);
    std::fflush(f_);
}

// ---- solution file ---------------------------------------------------------------------

NetCDFWriter::NetCDFWriter(const std::string &path, const HostGrid &grid,
                           const std::string &unit_x, const std::string &unit_y,
                           const std::string &unit_h, const std::string &unit_t, int save_every)
    : file_(path), grid_(grid), buf_((std::size_t)grid.Nx * grid.Ny) {
    const int dt = file_.add_dim("time", 0);
    const int dx = file_.add_dim("x", (std::size_t)grid.Nx);
    const int dy = file_.add_dim("y", (std::size_t)grid.Ny);
    v_time_ = file_.add_var("time", NcClassicFile::NC_DOUBLE, {dt});
    v_x_ = file_.add_var("x", NcClassicFile::NC_DOUBLE, {dx});
    v_y_ = file_.add_var("y", NcClassicFile::NC_DOUBLE, {dy});
    v_h_ = file_.add_var("h", NcClassicFile::NC_DOUBLE, {dt, dx, dy});
    v_hu_ = file_.add_var("hu", NcClassicFile::NC_DOUBLE, {dt, dx, dy});
    v_hv_ = file_.add_var("hv", NcClassicFile::NC_DOUBLE, {dt, dx, dy});
    v_B_ = file_.add_var("B", NcClassicFile::NC_DOUBLE, {dx, dy});
    if (!unit_x.empty()) file_.put_att(v_x_, "units", unit_x);
    if (!unit_y.empty()) file_.put_att(v_y_, "units", unit_y);
    if (!unit_h.empty()) {
        file_.put_att(v_h_, "units", unit_h);
        file_.put_att(v_B_, "units", unit_h);
    }
    if (!unit_t.empty()) file_.put_att(v_time_, "units", unit_t);
    file_.put_att(-1, "save_every", (int32_t)save_every);
}

NetCDFWriter::~NetCDFWriter() {
    try {
        if (file_.defining()) finish_header();
    } catch (...) {
    }
}

void NetCDFWriter::pack(const double *A, std::vector<double> &buf) const {
    for (int i = 0; i < grid_.Nx; ++i)
        for (int j = 0; j < grid_.Ny; ++j)
            buf[(std::size_t)i * grid_.Ny + j] = A[grid_.idx(i + grid_.nG, j + grid_.nG)];
}

void NetCDFWriter::finish_header() {
    file_.end_define();
    std::vector<double> x((std::size_t)grid_.Nx), y((std::size_t)grid_.Ny);
    for (int i = 0; i < grid_.Nx; ++i) x[(std::size_t)i] = grid_.x_center(i);
    for (int j = 0; j < grid_.Ny; ++j) y[(std::size_t)j] = grid_.y_center(j);
    file_.put_fixed(v_x_, x.data());
    file_.put_fixed(v_y_, y.data());
    if (have_B_) file_.put_fixed(v_B_, B_.data());
}

void NetCDFWriter::write_bathymetry(const double *B) {
    if (have_B_) return;
    B_.resize((std::size_t)grid_.Nx * grid_.Ny);
    pack(B, B_);
    have_B_ = true;
    if (!file_.defining()) file_.put_fixed(v_B_, B_.data());
}

void NetCDFWriter::write_snapshot(const double *h, const double *hu, const double *hv,
                                  double time, double dt, const std::string &backend,
                                  const std::string &riemann, const std::string &reconstruction,
                                  const std::string &time_integrator, const std::string &boundary,
                                  const std::string &bathymetry) {
    if (file_.defining()) {
        // the run's metadata comes with the first snapshot (netCDF_writer.cpp:105-136)
        if (!std::isnan(dt)) file_.put_att(-1, "dt", dt);
        if (!backend.empty()) file_.put_att(-1, "backend", backend);
        if (!riemann.empty()) file_.put_att(-1, "riemann_solver", riemann);
        if (!reconstruction.empty()) file_.put_att(-1, "reconstruction", reconstruction);
        if (!time_integrator.empty()) file_.put_att(-1, "time_integrator", time_integrator);
        if (!boundary.empty()) file_.put_att(-1, "boundary_condition", boundary);
        if (!bathymetry.empty()) file_.put_att(-1, "bathymetry", bathymetry);
        finish_header();
    }
    file_.put_record(v_time_, next_, &time);
    pack(h, buf_);
    file_.put_record(v_h_, next_, buf_.data());
    pack(hu, buf_);
    file_.put_record(v_hu_, next_, buf_.data());
    pack(hv, buf_);
    file_.put_record(v_hv_, next_, buf_.data());
    ++next_;
}

// ---- sanity-check file -----------------------------------------------------------------

SanityCheckNetCDFWriter::SanityCheckNetCDFWriter(
    const std::string &path, const std::string &time_unit, const std::string &h_unit,
    int save_every, double dt, const std::string &backend, const std::string &riemann,
    const std::string &reconstruction, const std::string &time_integrator,
    const std::string &boundary, const std::string &bathymetry)
    : file_(path) {
    const int rec = file_.add_dim("record", 0);
    v_step_ = file_.add_var("step", NcClassicFile::NC_INT, {rec});
    v_time_ = file_.add_var("time", NcClassicFile::NC_DOUBLE, {rec});
    v_err_ = file_.add_var("mass_rel_err", NcClassicFile::NC_DOUBLE, {rec});
    v_hmin_ = file_.add_var("h_min", NcClassicFile::NC_DOUBLE, {rec});
    file_.put_att(v_step_, "long_name", std::string("time_step_index"));
    file_.put_att(v_time_, "long_name", std::string("physical_time"));
    if (!time_unit.empty()) file_.put_att(v_time_, "units", time_unit);
    file_.put_att(v_err_, "long_name", std::string("relative_mass_error"));
    file_.put_att(v_err_, "units", std::string("1"));
    file_.put_att(v_hmin_, "long_name", std::string("minimum_water_height"));
    if (!h_unit.empty()) file_.put_att(v_hmin_, "units", h_unit);
    file_.put_att(-1, "save_every", (int32_t)save_every);
    file_.put_att(-1, "dt", dt);
    file_.put_att(-1, "backend", backend);
    file_.put_att(-1, "riemann_solver", riemann);
    file_.put_att(-1, "reconstruction", reconstruction);
    file_.put_att(-1, "time_integrator", time_integrator);
    file_.put_att(-1, "boundary_condition", boundary);
    file_.put_att(-1, "bathymetry", bathymetry);
    file_.put_att(-1, "mass_conservation_enabled", (int32_t)1);
    file_.put_att(-1, "positivity_enabled", (int32_t)1);
    file_.end_define();
}

void SanityCheckNetCDFWriter::write(std::size_t step, double time, double rel_err,
                                    double h_min) {
    const int32_t s = (int32_t)step;
    file_.put_record(v_step_, next_, &s);
    file_.put_record(v_time_, next_, &time);
    file_.put_record(v_err_, next_, &rel_err);
    file_.put_record(v_hmin_, next_, &h_min);
    ++next_;
}
