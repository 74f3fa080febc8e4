// netcdf_classic.hpp -- writers of the solver's output files in the NetCDF classic format
// (CDF-2, "64-bit offset"), implemented from the published format specification (no NetCDF
// library: none is installed).  netCDF, xarray and scipy.io.netcdf_file read the files.
//
// NetCDFWriter and SanityCheckNetCDFWriter produce the dimensions, variables and attributes
// of the reference's writers (include/io/netCDF_writer.cpp:30-143,
// include/io/sanity_checks_netcdf_writer.cpp:28-84); the one difference is the step counter
// of the sanity file, stored as a 32-bit int (classic NetCDF has no 64-bit integer type).
#pragma once

#include <cstdint>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "setup.hpp"

class NcClassicFile {
  public:
    enum Type : int32_t { NC_CHAR = 2, NC_INT = 4, NC_DOUBLE = 6 };

    explicit NcClassicFile(const std::string &path);  // creates parent directories
    ~NcClassicFile();
    NcClassicFile(const NcClassicFile &) = delete;
    NcClassicFile &operator=(const NcClassicFile &) = delete;

    // Definition phase.  A dimension of size 0 is the record (unlimited) dimension.
    int add_dim(const std::string &name, std::size_t size);
    int add_var(const std::string &name, Type type, const std::vector<int> &dims);
    void put_att(int var, const std::string &name, const std::string &text);  // var -1: global
    void put_att(int var, const std::string &name, int32_t value);
    void put_att(int var, const std::string &name, double value);
    bool defining() const { return !header_written_; }
    // Writes the header; afterwards only data may be written.
    void end_define();

    // Data phase (values in host order; converted to the file's big-endian order).
    void put_fixed(int var, const double *data);
    void put_record(int var, std::size_t rec, const double *data);
    void put_record(int var, std::size_t rec, const int32_t *data);
    std::size_t num_records() const { return numrecs_; }

  private:
    struct Att {
        std::string name;
        Type type;
        std::vector<unsigned char> bytes;  // big-endian values
        uint32_t nelems;
    };
    struct Dim {
        std::string name;
        std::size_t size;
    };
    struct Var {
        std::string name;
        Type type;
        std::vector<int> dims;
        std::vector<Att> atts;
        uint64_t vsize = 0, begin = 0;
        std::size_t nelem_slice = 0;  // elements of one record (record var) or all (fixed)
        bool record = false;
    };

    void write_at(uint64_t off, const void *p, std::size_t n);
    void update_numrecs(std::size_t rec);
    static std::size_t type_size(Type t) { return t == NC_CHAR ? 1 : (t == NC_INT ? 4 : 8); }

    std::FILE *f_ = nullptr;
    std::string path_;
    std::vector<Dim> dims_;
    std::vector<Var> vars_;
    std::vector<Att> gatts_;
    bool header_written_ = false;
    uint64_t recsize_ = 0;
    std::size_t numrecs_ = 0;
};

// The solution file (include/io/netCDF_writer.cpp): dims time (record), x, y; variables
// time(time), x(x), y(y), h/hu/hv(time, x, y), B(x, y); unit attributes; global save_every
// and, with the first snapshot, the run's metadata (dt, backend, schemes, bathymetry).
class NetCDFWriter {
  public:
    NetCDFWriter(const std::string &path, const HostGrid &grid, const std::string &unit_x,
                 const std::string &unit_y, const std::string &unit_h, const std::string &unit_t,
                 int save_every);
    ~NetCDFWriter();
    void write_bathymetry(const double *B);  // full array in the reference layout
    void write_snapshot(const double *h, const double *hu, const double *hv, double time,
                        double dt = std::numeric_limits<double>::quiet_NaN(),
                        const std::string &backend = "", const std::string &riemann = "",
                        const std::string &reconstruction = "",
                        const std::string &time_integrator = "",
                        const std::string &boundary = "", const std::string &bathymetry = "");

  private:
    void finish_header();
    void pack(const double *A, std::vector<double> &buf) const;

    NcClassicFile file_;
    HostGrid grid_;
    int v_time_, v_x_, v_y_, v_h_, v_hu_, v_hv_, v_B_;
    std::vector<double> buf_, B_;
    bool have_B_ = false;
    std::size_t next_ = 0;
};

// The sanity-check record file (include/io/sanity_checks_netcdf_writer.cpp).
class SanityCheckNetCDFWriter {
  public:
    SanityCheckNetCDFWriter(const std::string &path, const std::string &time_unit,
                            const std::string &h_unit, int save_every, double dt,
                            const std::string &backend, const std::string &riemann,
                            const std::string &reconstruction, const std::string &time_integrator,
                            const std::string &boundary, const std::string &bathymetry);
    void write(std::size_t step, double time, double rel_err, double h_min);

  private:
    NcClassicFile file_;
    int v_step_, v_time_, v_err_, v_hmin_;
    std::size_t next_ = 0;
};
