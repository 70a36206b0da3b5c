// Conformation files: the reader of redvar.f:13-648 (commands
// 'molecule : residue : variable : value', '&' fixes variables, '@' lines
// give global parameters) and the writer of outvar.f:12-121.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "state.hpp"

namespace smmp {

namespace {

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(' ');
  if (b == std::string::npos) return "";
  size_t e = s.find_last_not_of(' ');
  return s.substr(b, e - b + 1);
}

// extstr (redstr.f:15-86): text from ib up to the separator (or ie),
// trimmed; ib moves past the separator.
std::string extstr(char spr, size_t& ib, size_t ie, const std::string& str) {
  if (ib > ie) return "";
  std::string part = str.substr(ib, ie - ib + 1);
  size_t is = part.find(spr);
  std::string out;
  if (is == std::string::npos) {
    out = part;
    ib = ie + 1;
  } else if (is == 0) {
    ib = ib + 1;
    return "";
  } else {
    out = part.substr(0, is);
    ib = ib + is + 1;
  }
  return trim(out);
}

// iredin (redstr.f:140-178): unsigned integer of at most 9 digits.
bool iredin(const std::string& s_in, int& v) {
  std::string s = trim(s_in);
  if (s.empty() || s.size() > 9) return false;
  for (char c : s)
    if (c < '0' || c > '9') return false;
  v = std::atoi(s.c_str());
  return true;
}

// iredrl (redstr.f:180-228): real number, no comma, at most 6 decimals.
bool iredrl(const std::string& s_in, double& v) {
  std::string s = trim(s_in);
  if (s.empty() || s.find(',') != std::string::npos) return false;
  size_t ip = s.find('.');
  if (ip != std::string::npos) {
    if (s.size() > 17 || ip >= 11 || s.size() - 1 - ip > 6) return false;
  } else if (s.size() >= 11) {
    return false;
  }
  for (auto& c : s)
    if (c == 'd' || c == 'D') c = 'e';
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == '\0';
}

bool name_match(const std::string& have_in, const std::string& want, bool exact) {
  std::string have = trim(have_in);
  if (have.empty()) return false;
  size_t id = want.size();
  if (exact) return have.size() == id && have == want;
  return have.size() >= id && have.compare(0, id, want) == 0;
}

}  // namespace

void redvar_file(State& s, const std::string& varfile) {
  int ntlvr = s.nvars();
  if (ntlvr == 0) return;
  std::ifstream in(varfile);
  if (!in) throw Error("redvar> ERROR opening file to set variables: " + varfile);
  std::vector<char> forml(mxml + 1, 1), forrs(mxrs + 1, 1), forvr(ntlvr + 1, 1), stvr(ntlvr + 1, 0);
  std::vector<double> vlvrx(mxvr + 1, 0.0);
  for (int i = 1; i <= ntlvr; ++i) {
    s.fxvr(i) = 0;
    int it = s.ityvr(i);
    if (it == 3) vlvrx[i] = s.toat(s.iatvr(i));
    else if (it == 2) vlvrx[i] = s.baat(s.iatvr(i));
    else if (it == 1) vlvrx[i] = s.blat(s.iatvr(i));
    else throw Error("redvar> unknown variable type");
  }
  std::string line;
  int ncmd = 0;
  while (std::getline(in, line)) {
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\n')) line.pop_back();
    if (line.empty() || line.find('#') != std::string::npos) continue;
    size_t at = line.find('@');
    if (at != std::string::npos) {  // global parameters of a molecule
      std::string rest = line.substr(at + 1);
      int nml = 1;
      size_t colon = rest.find(':');
      if (colon != std::string::npos) {
        if (!iredin(rest.substr(0, colon), nml) || nml <= 0 || nml > s.ntlml)
          throw Error("redvar> Incorrect molecule number in line " + line);
        rest = rest.substr(colon + 1);
      }
      std::vector<double> rn;
      std::stringstream ss(rest);
      std::string tok;
      while (std::getline(ss, tok, ',')) {
        double v;
        if (!iredrl(tok, v)) throw Error("redvar> line with global parameters must contain 6 numbers: " + line);
        rn.push_back(v);
      }
      if (rn.size() != 6) throw Error("redvar> line with global parameters must contain 6 numbers: " + line);
      if (std::fabs(rn[3]) > 180.0 + 1e-6 || std::fabs(rn[4]) > 90.0 + 1e-6 || std::fabs(rn[5]) > 180.0 + 1e-6)
        throw Error("redvar> global angles out of range: " + line);
      for (int i = 1; i <= 3; ++i) s.gbpr(i, nml) = rn[i - 1];
      for (int i = 4; i <= 6; ++i) s.gbpr(i, nml) = rn[i - 1] * cdr;
      continue;
    }
    size_t ilb = 0, ile = line.size() - 1;
    while (ilb <= ile) {
      std::string cmd = extstr(';', ilb, ile, line);
      if (cmd.empty()) continue;
      if (++ncmd > 5000) throw Error("redvar> too many commands");
      bool fix = false;
      size_t ifx = cmd.find('&');
      if (ifx != std::string::npos) {
        fix = true;
        cmd[ifx] = ' ';
        cmd = trim(cmd);
        if (cmd.empty()) cmd = "*";
      }
      // fields separated by ':'; an empty field means 'all'
      std::vector<std::string> fld;
      size_t icb = 0, ice = cmd.size() - 1;
      while (icb <= ice) {
        if (fld.size() >= 4) throw Error("redvar> Cannot interpret command >" + cmd + "<");
        std::string f = extstr(':', icb, ice, cmd);
        fld.push_back(f.empty() ? "*" : f);
      }
      int nfld = (int)fld.size();
      for (int i = 0; i < nfld - 1; ++i) {
        int ifld = nfld - 1 - i;  // 3 molecules, 2 residues, 1 variables
        if (ifld == 3) for (int j = 1; j <= s.ntlml; ++j) forml[j] = 0;
        else if (ifld == 2) for (int j = 1; j <= s.ntlml; ++j) for (int k = s.irsml1(j); k <= s.irsml2(j); ++k) forrs[k] = 0;
        else if (ifld == 1) for (int j = 1; j <= ntlvr; ++j) forvr[j] = 0;
        size_t ifb = 0, ife = fld[i].size() - 1;
        int nide = 0;
        while (ifb <= ife) {
          if (++nide > 30) throw Error("redvar> too many identifiers in " + fld[i]);
          std::string ide = extstr(',', ifb, ife, fld[i]);
          auto bad = [&]() { return Error("redvar> Error in identifier >" + ide + "< of command >" + cmd + "<"); };
          if (ide.empty() || ide[0] == '*') {  // all
            for (int j = 1; j <= s.ntlml; ++j) {
              if (ifld == 3) forml[j] = 1;
              if (!forml[j]) continue;
              for (int k = s.irsml1(j); k <= s.irsml2(j); ++k) {
                if (ifld == 2) forrs[k] = 1;
                if (ifld == 1 && forrs[k])
                  for (int l = s.ivrrs1(k); l <= s.ivrrs1(k) + s.nvrrs(k) - 1; ++l) forvr[l] = 1;
              }
            }
            continue;
          }
          size_t ihy = ide.find('-');
          int inum;
          if (ihy == std::string::npos) {
            if (iredin(ide, inum)) {  // an index
              if (ifld == 3) {
                if (inum <= 0 || inum > s.ntlml) throw bad();
                forml[inum] = 1;
              }
              for (int j = 1; j <= s.ntlml; ++j) {
                if (!forml[j]) continue;
                if (ifld == 2) {
                  int k = inum + s.irsml1(j) - 1;
                  if (k < s.irsml1(j) || k > s.irsml2(j)) throw bad();
                  forrs[k] = 1;
                } else if (ifld == 1) {
                  for (int k = s.irsml1(j); k <= s.irsml2(j); ++k) {
                    if (!forrs[k]) continue;
                    int nfi = s.ivrrs1(k), l = inum + nfi - 1;
                    if (l < nfi || l > nfi + s.nvrrs(k) - 1) throw bad();
                    forvr[l] = 1;
                  }
                }
              }
            } else {  // a name, '*' at the end matches any ending
              bool exa = ide.back() != '*';
              std::string want = exa ? ide : ide.substr(0, ide.size() - 1);
              for (int j = 1; j <= s.ntlml; ++j) {
                if (ifld == 3) {
                  if (name_match(s.nmml.get(j), want, exa)) forml[j] = 1;
                  continue;
                }
                if (!forml[j]) continue;
                for (int k = s.irsml1(j); k <= s.irsml2(j); ++k) {
                  if (ifld == 2) {
                    if (name_match(s.seq.get(k), want, exa)) forrs[k] = 1;
                  } else if (ifld == 1 && forrs[k]) {
                    for (int l = s.ivrrs1(k); l <= s.ivrrs1(k) + s.nvrrs(k) - 1; ++l)
                      if (name_match(s.nmvr.get(l), want, exa)) forvr[l] = 1;
                  }
                }
              }
            }
          } else {  // a zone n1-n2, -n2 or n1-
            if (ide.size() == 1) throw bad();
            int ibz = 1, iez = 1000000;
            if (ihy > 0 && (!iredin(ide.substr(0, ihy), ibz) || ibz <= 0)) throw bad();
            if (ihy + 1 < ide.size() && (!iredin(ide.substr(ihy + 1), iez) || iez <= 0 || iez < ibz)) throw bad();
            for (int j = 1; j <= s.ntlml; ++j) {
              if (ifld == 3) {
                if (j >= ibz && j <= std::min(iez, (int)s.ntlml)) forml[j] = 1;
                continue;
              }
              if (!forml[j]) continue;
              if (ifld == 2) {
                int kbz = s.irsml1(j) + ibz - 1, kez = std::min(s.irsml1(j) + iez - 1, s.irsml2(j));
                for (int k = kbz; k <= kez; ++k) forrs[k] = 1;
              } else if (ifld == 1) {
                for (int k = s.irsml1(j); k <= s.irsml2(j); ++k) {
                  int kv = s.nvrrs(k);
                  if (!forrs[k] || kv <= 0) continue;
                  int ll = s.ivrrs1(k), lbz = ll + ibz - 1, lez = (iez > kv) ? ll + kv - 1 : ll + iez - 1;
                  for (int l = lbz; l <= lez; ++l) forvr[l] = 1;
                }
              }
            }
          }
        }
      }
      double val;
      if (!iredrl(fld[nfld - 1], val)) throw Error("redvar> Cannot read value from >" + cmd + "<");
      for (int i = 1; i <= ntlvr; ++i)
        if (forvr[i]) {
          vlvrx[i] = val;
          s.fxvr(i) = fix ? 1 : 0;
          stvr[i] = 1;
        }
    }
  }
  // apply: degrees -> radians for the variables that were set, rebuild
  int iv = 0;
  for (int i = 1; i <= s.ntlml; ++i) {
    bool global = false;
    for (int j = 1; j <= 6; ++j)
      if (s.gbpr(j, i) != 0.0) global = true;
    if (global) setvar(s, i, vlvrx.data() + 1);
    if (s.nvrml(i) > 0) {
      bool did = false;
      for (int j = s.irsml1(i); j <= s.irsml2(i); ++j)
        for (int k = s.ivrrs1(j); k <= s.ivrrs1(j) + s.nvrrs(j) - 1; ++k) {
          ++iv;
          if (stvr[iv]) {
            did = true;
            int ity = s.ityvr(iv);
            if (ity == 3 || ity == 2) vlvrx[iv] = vlvrx[iv] * cdr;
          }
        }
      if (did) setvar(s, i, vlvrx.data() + 1);
    }
  }
}

void write_var_file(const State& s, int nml, const std::string& path) {
  if (nml < 0 || nml > s.ntlml) throw Error("outvar> No such molecule");
  FILE* f = path.empty() ? stdout : std::fopen(path.c_str(), "w");
  if (!f) throw Error("outvar> cannot open " + path);
  int im1 = nml > 0 ? nml : 1, im2 = nml > 0 ? nml : s.ntlml;
  for (int iml = im1; iml <= im2; ++iml) {
    char mlfd[16];
    std::snprintf(mlfd, sizeof mlfd, "%4d : ", iml);
    bool global = false;
    for (int j = 1; j <= 6; ++j)
      if (s.gbpr(j, iml) != 0.0) global = true;
    if (global) {
      std::fprintf(f, " @ %s ", mlfd);
      for (int i = 1; i <= 6; ++i) {
        double v = i <= 3 ? s.gbpr(i, iml) : s.gbpr(i, iml) * crd;
        std::fprintf(f, i < 6 ? "%.6f," : "%.6f\n", v);
      }
    }
    int ifivr = s.ivrml1(iml), is = s.irsml1(iml) - 1;
    for (int i = ifivr; i <= ifivr + s.nvrml(iml) - 1; ++i) {
      std::string nm = s.nmvr.get(i);
      if (!s.fxvr(i))
        std::fprintf(f, "   %s%3d : %s : %10.3f\n", mlfd, nursvr(s, i) - is, nm.c_str(), s.vlvr(i) * crd);
      else if (s.ityvr(i) == 3)
        std::fprintf(f, "   %s%3d : %s : %10.3f  &\n", mlfd, nursvr(s, i) - is, nm.c_str(), s.vlvr(i) * crd);
    }
  }
  if (f != stdout) std::fclose(f);
}

}  // namespace smmp
