// xerus's datafile format for TTTensor / TTOperator, byte compatible in both directions.
//
// File = header line "Xerus xerus::TTNetwork<false|true> datafile." + "Format: Binary|TSV" line + payload
// (misc/fileIO.h:102-163). The payload is the nested stream writers of TTNetwork (ttNetwork.cpp:1454-1470:
// version, canonicalized, corePosition), TensorNetwork (tensorNetwork.cpp:1429-1469: version, dimensions,
// external links, node links, node tensors) and Tensor (tensor.cpp:1781-1806: version, dimensions,
// representation, values). Scalars are raw little-endian size_t / bool / double in the binary format and
// tab-terminated decimal text (16 significant digits) in TSV, where the TensorNetwork writer adds its line breaks.
//
// The network written for a TT of d components (N = 1 tensor, 2 operator) is the one xerus builds
// (ttNetwork.cpp:57-108): node 0 and node d+1 are the virtual boundary tensors ones({1}), node k+1 holds
// component k with links [left rank, mode(s), right rank]. Cores are written dense; the reader also accepts
// sparse tensors (tensor.cpp:1828-1842), as xerus writes them for dirac-initialised components.
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>

#include "tt.hpp"

namespace xtt {

namespace {

constexpr size_t kNone = ~size_t(0);  // Link::other of an external index (size_t(-1))

const char* class_name(bool op) { return op ? "xerus::TTNetwork<true>" : "xerus::TTNetwork<false>"; }

struct Writer {
	std::ostream& os;
	bool tsv;
	void precision() {
		if (tsv) os << std::setprecision(std::numeric_limits<double>::digits10 + 1);
	}
	void sz(size_t v) {
		if (tsv) os << v << '\t';
		else os.write(reinterpret_cast<const char*>(&v), sizeof v);
	}
	void flag(bool v) {
		if (tsv) os << v << '\t';
		else os.write(reinterpret_cast<const char*>(&v), sizeof v);
	}
	void value(double v) {
		if (tsv) os << v << '\t';
		else os.write(reinterpret_cast<const char*>(&v), sizeof v);
	}
	void dims(const dims_t& d) {
		sz(d.size());
		for (size_t x : d) sz(x);
	}
	void text(const char* s) {
		if (tsv) os << s;
	}
	void link(bool external, size_t other, size_t pos, size_t dim) {
		flag(external);
		sz(other);
		sz(pos);
		sz(dim);
	}
	void dense_tensor(const dims_t& d, const std::vector<double>& v) {  // tensor.cpp:1781-1806, dense branch
		precision();
		sz(1);
		dims(d);
		sz(1);
		for (double x : v) value(x);
	}
};

struct Reader {
	std::istream& is;
	bool tsv;
	const std::string& file;
	void fail(const std::string& what) const { throw Error(ERR_INVALID, "load_from_file(" + file + "): " + what); }
	void check() const {
		if (!is) fail("unexpected end of stream");
	}
	size_t sz() {
		size_t v = 0;
		if (tsv) is >> v;
		else is.read(reinterpret_cast<char*>(&v), sizeof v);
		check();
		return v;
	}
	bool flag() {
		bool v = false;
		if (tsv) is >> v;
		else is.read(reinterpret_cast<char*>(&v), sizeof v);
		check();
		return v;
	}
	double value() {
		double v = 0;
		if (tsv) is >> v;
		else is.read(reinterpret_cast<char*>(&v), sizeof v);
		check();
		return v;
	}
	dims_t dims() {
		const size_t n = sz();
		if (n > 4096) fail("implausible tensor degree");
		dims_t d(n);
		for (size_t& x : d) x = sz();
		return d;
	}
	void version(const char* what) {
		const size_t v = sz();
		if (v != 1) fail(std::string("unknown ") + what + " stream version " + std::to_string(v));
	}
	// tensor.cpp:1810-1843: dense values, or (position, value) pairs of a sparse tensor
	std::vector<double> tensor(dims_t& d) {
		version("Tensor");
		d = dims();
		size_t count = 1;
		for (size_t x : d) {
			if (x == 0 || count > (size_t(1) << 34) / x) fail("tensor too large or with a zero dimension");
			count *= x;
		}
		std::vector<double> v(count, 0.0);
		const size_t rep = sz();
		if (rep == 1) {
			if (tsv) {
				for (double& x : v) x = value();
			} else {
				is.read(reinterpret_cast<char*>(v.data()), std::streamsize(count * sizeof(double)));
				check();
			}
		} else if (rep == 2) {
			const size_t nnz = sz();
			for (size_t i = 0; i < nnz; ++i) {
				const size_t pos = sz();
				const double val = value();
				if (pos >= count) fail("sparse entry outside the tensor");
				v[pos] = val;
			}
		} else {
			fail("unknown tensor representation " + std::to_string(rep));
		}
		return v;
	}
};

void write_tt(Writer& w, const TT& x) {
	const size_t d = x.order(), N = x.op ? 2 : 1;
	const dims_t r = x.ranks();  // d-1 inner ranks
	auto rank = [&](size_t i) { return i == 0 || i == d ? size_t(1) : r[i - 1]; };
	// TTNetwork (ttNetwork.cpp:1454-1470)
	w.precision();
	w.sz(1);
	w.flag(x.canonicalized);
	w.sz(x.corePosition);
	// TensorNetwork (tensorNetwork.cpp:1429-1469)
	w.precision();
	w.sz(1);
	w.dims(x.dims);
	w.text("\n");
	for (size_t i = 0; i < d; ++i) {  // external links: mode i of component i at index position 1 (and 2)
		w.sz(i + 1);
		w.sz(1);
		w.sz(x.dims[i]);
	}
	if (x.op)
		for (size_t i = 0; i < d; ++i) {
			w.sz(i + 1);
			w.sz(2);
			w.sz(x.dims[d + i]);
		}
	w.text("\n\n");
	w.sz(d + 2);
	w.text("\n");
	w.sz(1);
	w.link(false, 1, 0, 1);
	for (size_t i = 0; i < d; ++i) {
		w.sz(N + 2);
		w.link(false, i, i == 0 ? 0 : N + 1, rank(i));
		w.link(true, kNone, i, x.dims[i]);
		if (x.op) w.link(true, kNone, i + d, x.dims[d + i]);
		w.link(false, i + 2, 0, rank(i + 1));
	}
	w.sz(1);
	w.link(false, d, N + 1, 1);
	w.text("\n");
	const std::vector<double> one{1.0};
	w.dense_tensor({1}, one);
	w.text("\n");
	for (size_t i = 0; i < d; ++i) {
		std::vector<double> h(x.cores[i].count());
		dt_download(x.cores[i], h.data());
		w.dense_tensor(x.cores[i].dims, h);
		w.text("\n");
	}
	w.dense_tensor({1}, one);
	w.text("\n");
}

TT read_tt(Reader& rd, bool op) {
	rd.version("TTNetwork");
	const bool canonicalized = rd.flag();
	const size_t corePosition = rd.sz();
	rd.version("TensorNetwork");
	const dims_t dims = rd.dims();
	const size_t N = op ? 2 : 1;
	if (dims.empty() || dims.size() % N != 0) rd.fail("degree-0 or odd-degree TT networks are not supported");
	const size_t d = dims.size() / N;
	for (size_t n = 0; n < dims.size(); ++n) {  // external links (ttNetwork.cpp:299-302)
		const size_t other = rd.sz(), pos = rd.sz(), dim = rd.sz();
		if (other != n % d + 1 || pos != 1 + n / d || dim != dims[n]) rd.fail("external link " + std::to_string(n) + " is not a TT link");
	}
	const size_t nodes = rd.sz();
	if (nodes != d + 2) rd.fail("wrong number of nodes for a TT network");
	std::vector<std::vector<size_t>> linkDims(nodes);
	for (size_t k = 0; k < nodes; ++k) {
		const size_t deg = rd.sz();
		const size_t want = (k == 0 || k == nodes - 1) ? 1 : N + 2;
		if (deg != want) rd.fail("node " + std::to_string(k) + " has the wrong degree for a TT network");
		for (size_t l = 0; l < deg; ++l) {
			const bool external = rd.flag();
			const size_t other = rd.sz(), pos = rd.sz(), dim = rd.sz();
			bool ok;
			if (k == 0) ok = !external && other == 1 && pos == 0 && dim == 1;
			else if (k == nodes - 1) ok = !external && other == d && pos == N + 1 && dim == 1;
			else if (l == 0) ok = !external && other == k - 1 && pos == (k == 1 ? 0 : N + 1);
			else if (l == N + 1) ok = !external && other == k + 1 && pos == 0;
			else ok = external && other == kNone && pos == (k - 1) + (l - 1) * d && dim == dims[pos];
			if (!ok) rd.fail("node " + std::to_string(k) + " link " + std::to_string(l) + " is not a TT link");
			linkDims[k].push_back(dim);
		}
	}
	TT x;
	x.op = op;
	x.dims = dims;
	x.cores.resize(d);
	for (size_t k = 0; k < nodes; ++k) {
		dims_t td;
		std::vector<double> v = rd.tensor(td);
		if (td != linkDims[k]) rd.fail("tensor of node " + std::to_string(k) + " does not match its links");
		if (k == 0 || k == nodes - 1) {
			if (v[0] != 1.0) rd.fail("a virtual boundary node is not 1.0");
			continue;
		}
		if (k == 1 && td[0] != 1) rd.fail("the first component's left rank is not 1");
		if (k == d && td.back() != 1) rd.fail("the last component's right rank is not 1");
		if (k > 1 && td[0] != linkDims[k - 1].back()) rd.fail("rank mismatch between components");
		x.cores[k - 1] = dt_upload(td, v.data());
	}
	if (canonicalized && corePosition >= d) rd.fail("invalid core position");
	x.canonicalized = canonicalized;
	x.corePosition = corePosition;
	return x;
}

}  // namespace

void tt_save(const TT& x, const std::string& filename, int format) {  // misc::save_to_file, fileIO.h:105-121
	XTT_REQUIRE(format == 0 || format == 1, "save_to_file: format must be BINARY (0) or TSV (1)");
	XTT_REQUIRE(x.order() > 0, "save_to_file: degree-0 TT networks are not supported");
	const bool tsv = format == 1;
	std::ofstream out(filename, tsv ? std::ofstream::out : (std::ofstream::out | std::ofstream::binary));
	if (!out) throw Error(ERR_INVALID, "save_to_file: cannot open " + filename);
	out << "Xerus " << class_name(x.op) << " datafile.\nFormat: " << (tsv ? "TSV" : "Binary") << "\n";
	Writer w{out, tsv};
	write_tt(w, x);
	out.flush();
	if (!out) throw Error(ERR_INVALID, "save_to_file: write to " + filename + " failed");
}

TT tt_load(const std::string& filename) {  // misc::load_from_file, fileIO.h:125-160
	std::ifstream in(filename, std::ifstream::in);
	if (!in) throw Error(ERR_INVALID, "load_from_file: cannot open " + filename);
	std::string first;
	std::getline(in, first);
	if (!in) throw Error(ERR_INVALID, "load_from_file(" + filename + "): unexpected end of stream");
	bool op;
	if (first == std::string("Xerus ") + class_name(false) + " datafile.") op = false;
	else if (first == std::string("Xerus ") + class_name(true) + " datafile.") op = true;
	else throw Error(ERR_INVALID, "load_from_file(" + filename + "): not a TTTensor / TTOperator datafile: " + first);
	std::string qual, val;
	in >> qual >> val;
	if (qual != "Format:") throw Error(ERR_INVALID, "load_from_file(" + filename + "): invalid syntax " + qual);
	bool tsv;
	if (val == "TSV") {
		tsv = true;
	} else if (val == "Binary") {
		tsv = false;
		const std::streamoff pos = in.tellg();  // reopen in binary mode past the line break
		in.close();
		in.open(filename, std::ifstream::in | std::ifstream::binary);
		in.seekg(pos + 1);
	} else {
		throw Error(ERR_INVALID, "load_from_file(" + filename + "): invalid format " + val);
	}
	Reader rd{in, tsv, filename};
	return read_tt(rd, op);
}

}  // namespace xtt
