// libnbl_hdf5.so: reading and writing Nebula .mat material files with the
// HDF5 C library, kept out of libnebula.so so that the tracking library does
// not depend on HDF5.
//
// File conventions (io/hdf5_file.cpp of the reference, paths relative to its
// source/): quantities are attributes of compound type {double value;
// variable-length string unit}; datasets carry a "units" string attribute and
// one HDF5 dimension scale per axis, each with its own "units" attribute.
#include <hdf5.h>
#include <hdf5_hl.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/nebula.h"

namespace {

// ------------------------------------------------------------ unit parser --
// Units of common/units/unit_system.h:31-46 in internal units (eV, nm, ns,
// K, e), parsed as in common/units/unit_parser.h:46-92: "a*b/c^2".
double parse_unit(const std::string& text)
{
	static const std::map<std::string, double> units = {
		{ "eV", 1.0 }, { "nm", 1.0 }, { "K", 1.0 }, { "s", 1e9 }, { "m", 1e9 },
		{ "cm", 1e7 }, { "g", 6.2415e15 }, { "C", 6.2415e18 }, { "radian", 1.0 } };
	double value = 1.0;
	size_t i = 0;
	const size_t n = text.size();
	auto is_new = [](char c) { return c == '*' || c == '/'; };
	while (i < n) {
		int power = (text[i] == '/') ? -1 : 1;
		if (is_new(text[i]))
			++i;
		std::string name;
		for (; i < n && text[i] != '^' && !is_new(text[i]); ++i)
			name.push_back(text[i]);
		auto it = units.find(name);
		if (it == units.end())
			throw std::runtime_error("Unknown unit '" + name + "'.");
		if (i < n && text[i] == '^') {
			std::string p;
			for (++i; i < n && !is_new(text[i]); ++i)
				p.push_back(text[i]);
			power *= std::stoi(p);
		}
		value *= std::pow(it->second, power);
	}
	return value;
}

hid_t vlen_string_type()
{
	hid_t t = H5Tcopy(H5T_C_S1);
	H5Tset_size(t, H5T_VARIABLE);
	return t;
}

std::string read_string_attr(hid_t obj, const char* name)
{
	hid_t a = H5Aopen(obj, name, H5P_DEFAULT);
	if (a < 0)
		throw std::runtime_error(std::string("Could not read string attribute '") + name + "'");
	hid_t st = vlen_string_type();
	char* buf = nullptr;
	H5Aread(a, st, &buf);
	std::string s = buf ? buf : "";
	if (buf) H5free_memory(buf);
	H5Tclose(st);
	H5Aclose(a);
	return s;
}

struct QuantityC { double value; char* unit; };

hid_t quantity_type()
{
	hid_t st = vlen_string_type();
	hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(QuantityC));
	H5Tinsert(t, "value", HOFFSET(QuantityC, value), H5T_NATIVE_DOUBLE);
	H5Tinsert(t, "unit", HOFFSET(QuantityC, unit), st);
	H5Tclose(st);
	return t;
}

// hdf5_file::get_property_quantity (io/hdf5_file.cpp:146-188).
bool read_quantity(hid_t file, const char* name, double& out)
{
	hid_t a = H5Aopen(file, name, H5P_DEFAULT);
	if (a < 0)
		return false;
	hid_t t = quantity_type();
	QuantityC q{ 0.0, nullptr };
	H5Aread(a, t, &q);
	const std::string unit = q.unit ? q.unit : "";
	if (q.unit) H5free_memory(q.unit);
	H5Tclose(t);
	H5Aclose(a);
	out = q.value * parse_unit(unit);
	return true;
}

struct Storage {
	std::vector<std::vector<float>> tables;
	std::vector<double> shells;
};

// Per axis: first and last value of the attached dimension scale (times its
// unit), as get_lin_dimscale / get_log_dimscale read them (:291-359).
void read_axes(hid_t ds, int ndim, nbl_raw_dataset& d)
{
	for (int a = 0; a < ndim; ++a) {
		struct It { std::vector<double> v; double unit = 1; bool found = false; } it;
		H5DSiterate_scales(ds, (unsigned)a, nullptr,
			[](hid_t, unsigned, hid_t sid, void* p) -> herr_t {
				auto* it = (It*)p;
				it->unit = parse_unit(read_string_attr(sid, "units"));
				hid_t sp = H5Dget_space(sid);
				hsize_t n = 0;
				H5Sget_simple_extent_dims(sp, &n, nullptr);
				H5Sclose(sp);
				it->v.resize(n);
				H5Dread(sid, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, it->v.data());
				it->found = true;
				return 1;
			}, &it);
		d.ax_present[a] = 0;
		d.ax_front[a] = 0;
		d.ax_back[a] = 1;
		if (it.found) {
			if ((int)it.v.size() != d.dims[a])
				throw std::runtime_error("Dimension scale has unexpected size.");
			d.ax_present[a] = 1;
			d.ax_front[a] = it.v.front() * it.unit;
			d.ax_back[a] = it.v.back() * it.unit;
		}
	}
}

// hdf5_file::fill_table{1,2,3}D + get_unit (:229-289, :361-375).
void read_dataset(hid_t file, const char* path, int ndim, nbl_raw_dataset& d, Storage& st)
{
	hid_t ds = H5Dopen(file, path, H5P_DEFAULT);
	if (ds < 0)
		throw std::runtime_error(std::string("Could not read table '") + path + "'");
	hid_t sp = H5Dget_space(ds);
	if (H5Sget_simple_extent_ndims(sp) != ndim) {
		H5Sclose(sp);
		H5Dclose(ds);
		throw std::runtime_error("Dataspace has unexpected dimension");
	}
	hsize_t dims[3] = { 0, 0, 0 };
	H5Sget_simple_extent_dims(sp, dims, nullptr);
	H5Sclose(sp);
	size_t count = 1;
	std::memset(&d, 0, sizeof(d));
	d.ndim = ndim;
	for (int a = 0; a < ndim; ++a) {
		d.dims[a] = (int32_t)dims[a];
		count *= dims[a];
	}
	st.tables.emplace_back(count);
	H5Dread(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, st.tables.back().data());
	d.data = st.tables.back().data();
	d.unit = parse_unit(read_string_attr(ds, "units"));
	read_axes(ds, ndim, d);
	H5Dclose(ds);
}

bool exists(hid_t file, const char* path) { return H5LTpath_valid(file, path, 0) == 1; }

// ------------------------------------------------------------------ writer --
struct Writer {
	hid_t file;
	int scale_counter = 0;
};

void write_string_attr(hid_t obj, const char* name, const char* value)
{
	hid_t st = vlen_string_type();
	hid_t sp = H5Screate(H5S_SCALAR);
	hid_t a = H5Acreate2(obj, name, st, sp, H5P_DEFAULT, H5P_DEFAULT);
	H5Awrite(a, st, &value);
	H5Aclose(a);
	H5Sclose(sp);
	H5Tclose(st);
}

hid_t ensure_groups(hid_t file, const std::string& path)
{
	// create intermediate groups of "/a/b/name"
	size_t pos = 1;
	while (true) {
		const size_t s = path.find('/', pos);
		if (s == std::string::npos)
			break;
		const std::string g = path.substr(0, s);
		if (!exists(file, g.c_str())) {
			hid_t gid = H5Gcreate2(file, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
			H5Gclose(gid);
		}
		pos = s + 1;
	}
	return 0;
}

} // namespace

extern "C" {

// Read a .mat file as the reference's material constructor does
// (core/material.inl:13-18 + the create() of each mechanism): attributes
// barrier, fermi, band_gap (optional, default -1 eV), phonon_loss,
// effective_A; tables of /elastic and of the inelastic model.
int nbl_h5_read_material(const char* path, int32_t inel_model, nbl_raw_material* out,
	void** storage, char* msg, size_t msg_len)
{
	H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
	auto* st = new Storage();
	hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0) {
		std::snprintf(msg, msg_len, "Unable to open file %s", path);
		delete st;
		return NBL_ERR_IO;
	}
	try {
		std::memset(out, 0, sizeof(*out));
		out->inel_model = inel_model;
		auto need = [&](const char* n, double& v) {
			if (!read_quantity(f, n, v))
				throw std::runtime_error(std::string("Could not read quantity property '") + n + "'");
		};
		need("barrier", out->barrier);
		need("fermi", out->fermi);
		need("phonon_loss", out->phonon_loss);
		need("effective_A", out->effective_A);
		out->has_band_gap = read_quantity(f, "band_gap", out->band_gap) ? 1 : 0;
		if (!exists(f, "/elastic"))
			throw std::runtime_error(std::string("Kieft elastic model not found in material file ") + path);
		read_dataset(f, "/elastic/imfp", 1, out->el_imfp, *st);
		read_dataset(f, "/elastic/costheta_icdf", 2, out->el_icdf, *st);
		if (inel_model == NBL_INEL_PENN) {
			if (!exists(f, "/full_penn"))
				throw std::runtime_error(std::string("Full Penn model not found in material file ") + path);
			read_dataset(f, "/full_penn/imfp", 1, out->in_imfp, *st);
			read_dataset(f, "/full_penn/omega_icdf", 2, out->in_icdf, *st);
			read_dataset(f, "/full_penn/q_icdf", 3, out->pn_q, *st);
			read_dataset(f, "/ionization/binding_icdf", 3, out->pn_ion, *st);
		} else {
			if (!exists(f, "/inelastic_kieft"))
				throw std::runtime_error(std::string("Kieft inelastic model not found in material file ") + path);
			read_dataset(f, "/inelastic_kieft/imfp", 1, out->in_imfp, *st);
			read_dataset(f, "/inelastic_kieft/w0_icdf", 2, out->in_icdf, *st);
			read_dataset(f, "/ionization/binding_icdf", 2, out->ki_ion, *st);
			nbl_raw_dataset sh;
			read_dataset(f, "/ionization/outer_shells", 1, sh, *st);
			for (int i = 0; i < sh.dims[0]; ++i)
				st->shells.push_back((double)sh.data[i]);
			out->outer_shells = st->shells.data();
			out->n_outer_shells = (int32_t)st->shells.size();
			out->outer_shells_unit = sh.unit;
		}
	} catch (const std::exception& e) {
		std::snprintf(msg, msg_len, "%s", e.what());
		H5Fclose(f);
		delete st;
		return NBL_ERR_FORMAT;
	}
	H5Fclose(f);
	*storage = st;
	return NBL_OK;
}

void nbl_h5_free_material(void* storage) { delete (Storage*)storage; }

// Writer used to produce .mat files from synthetic tables.
int nbl_h5_create(const char* path, void** handle)
{
	H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
	hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
	if (f < 0)
		return NBL_ERR_IO;
	*handle = new Writer{ f };
	return NBL_OK;
}

int nbl_h5_write_quantity(void* handle, const char* name, double value, const char* unit)
{
	auto* w = (Writer*)handle;
	hid_t t = quantity_type();
	hid_t sp = H5Screate(H5S_SCALAR);
	hid_t a = H5Acreate2(w->file, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
	QuantityC q{ value, const_cast<char*>(unit) };
	const herr_t e = H5Awrite(a, t, &q);
	H5Aclose(a);
	H5Sclose(sp);
	H5Tclose(t);
	return e < 0 ? NBL_ERR_IO : NBL_OK;
}

// Dataset of float32 data with a units attribute and, per axis, an optional
// float64 dimension scale (scales[a] may be NULL) with its own units.
int nbl_h5_write_dataset(void* handle, const char* path, int32_t ndim, const int64_t* dims,
	const float* data, const char* unit, const double* const* scales, const char* const* scale_units)
{
	auto* w = (Writer*)handle;
	ensure_groups(w->file, path);
	hsize_t hd[3];
	for (int a = 0; a < ndim; ++a)
		hd[a] = (hsize_t)dims[a];
	hid_t sp = H5Screate_simple(ndim, hd, nullptr);
	hid_t ds = H5Dcreate2(w->file, path, H5T_IEEE_F32LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	if (ds < 0) {
		H5Sclose(sp);
		return NBL_ERR_IO;
	}
	H5Dwrite(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
	write_string_attr(ds, "units", unit);
	for (int a = 0; a < ndim; ++a) {
		if (!scales || !scales[a])
			continue;
		const std::string sname = std::string("/_scales/s") + std::to_string(w->scale_counter++);
		ensure_groups(w->file, sname);
		hsize_t n = hd[a];
		hid_t ssp = H5Screate_simple(1, &n, nullptr);
		hid_t sds = H5Dcreate2(w->file, sname.c_str(), H5T_IEEE_F64LE, ssp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
		H5Dwrite(sds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, scales[a]);
		write_string_attr(sds, "units", scale_units && scale_units[a] ? scale_units[a] : "");
		H5DSset_scale(sds, nullptr);
		H5DSattach_scale(ds, sds, (unsigned)a);
		H5Dclose(sds);
		H5Sclose(ssp);
	}
	H5Dclose(ds);
	H5Sclose(sp);
	return NBL_OK;
}

int nbl_h5_close(void* handle)
{
	auto* w = (Writer*)handle;
	H5Fclose(w->file);
	delete w;
	return NBL_OK;
}

} // extern "C"
