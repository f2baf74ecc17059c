// ng_json.h -- minimal JSON reader for the dictionary arguments of the C ABI.
//
// NEST passes status, conn_spec and syn_spec dictionaries into its kernel API
// (nestkernel/nest.h: create, connect, set_kernel_status); across the C ABI they
// travel as JSON text. Numbers, strings, booleans, null, arrays and objects.
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nestgpu
{

struct Json
{
  enum Kind
  {
    NUL,
    BOOL,
    NUM,
    STR,
    ARR,
    OBJ
  } kind = NUL;
  bool b = false;
  double num = 0;
  bool is_int = false;
  std::string str;
  std::vector< Json > arr;
  std::vector< std::pair< std::string, Json > > obj;  // insertion order kept

  bool
  has( const std::string& k ) const
  {
    if ( kind != OBJ )
      return false;
    for ( const auto& kv : obj )
      if ( kv.first == k )
        return true;
    return false;
  }
  const Json&
  at( const std::string& k ) const
  {
    for ( const auto& kv : obj )
      if ( kv.first == k )
        return kv.second;
    throw std::runtime_error( "missing key '" + k + "'" );
  }
  double
  number() const
  {
    if ( kind == NUM )
      return num;
    if ( kind == BOOL )
      return b ? 1.0 : 0.0;
    throw std::runtime_error( "expected a number" );
  }
  const std::string&
  string() const
  {
    if ( kind != STR )
      throw std::runtime_error( "expected a string" );
    return str;
  }

  static Json
  parse( const char* text )
  {
    if ( text == nullptr or *text == 0 )
    {
      Json j;
      j.kind = OBJ;
      return j;
    }
    const char* p = text;
    Json j = parse_value( p );
    skip( p );
    if ( *p )
      throw std::runtime_error( std::string( "trailing characters in JSON: " ) + p );
    return j;
  }

private:
  static void
  skip( const char*& p )
  {
    while ( *p == ' ' or *p == '\n' or *p == '\t' or *p == '\r' )
      ++p;
  }
  static std::string
  parse_string( const char*& p )
  {
    if ( *p != '"' )
      throw std::runtime_error( "expected '\"' in JSON" );
    ++p;
    std::string s;
    while ( *p and *p != '"' )
    {
      if ( *p == '\\' )
      {
        ++p;
        switch ( *p )
        {
        case 'n':
          s += '\n';
          break;
        case 't':
          s += '\t';
          break;
        case 'u':
          // only ASCII escapes are expected in keys and model names
          s += static_cast< char >( std::strtol( std::string( p + 1, 4 ).c_str(), nullptr, 16 ) );
          p += 4;
          break;
        default:
          s += *p;
        }
        ++p;
      }
      else
        s += *p++;
    }
    if ( *p != '"' )
      throw std::runtime_error( "unterminated JSON string" );
    ++p;
    return s;
  }
  static Json
  parse_value( const char*& p )
  {
    skip( p );
    Json j;
    if ( *p == '{' )
    {
      j.kind = OBJ;
      ++p;
      skip( p );
      if ( *p == '}' )
      {
        ++p;
        return j;
      }
      for ( ;; )
      {
        skip( p );
        std::string k = parse_string( p );
        skip( p );
        if ( *p != ':' )
          throw std::runtime_error( "expected ':' in JSON" );
        ++p;
        j.obj.emplace_back( k, parse_value( p ) );
        skip( p );
        if ( *p == ',' )
        {
          ++p;
          continue;
        }
        if ( *p == '}' )
        {
          ++p;
          return j;
        }
        throw std::runtime_error( "expected ',' or '}' in JSON" );
      }
    }
    if ( *p == '[' )
    {
      j.kind = ARR;
      ++p;
      skip( p );
      if ( *p == ']' )
      {
        ++p;
        return j;
      }
      for ( ;; )
      {
        j.arr.push_back( parse_value( p ) );
        skip( p );
        if ( *p == ',' )
        {
          ++p;
          continue;
        }
        if ( *p == ']' )
        {
          ++p;
          return j;
        }
        throw std::runtime_error( "expected ',' or ']' in JSON" );
      }
    }
    if ( *p == '"' )
    {
      j.kind = STR;
      j.str = parse_string( p );
      return j;
    }
    if ( std::string( p, 4 ) == "true" )
    {
      j.kind = BOOL;
      j.b = true;
      p += 4;
      return j;
    }
    if ( std::string( p, 5 ) == "false" )
    {
      j.kind = BOOL;
      p += 5;
      return j;
    }
    if ( std::string( p, 4 ) == "null" )
    {
      p += 4;
      return j;
    }
    char* end = nullptr;
    j.num = std::strtod( p, &end );
    if ( end == p )
      throw std::runtime_error( std::string( "bad JSON value near: " ) + std::string( p, 20 ) );
    j.kind = NUM;
    j.is_int = true;
    for ( const char* q = p; q < end; ++q )
      if ( *q == '.' or *q == 'e' or *q == 'E' )
        j.is_int = false;
    p = end;
    return j;
  }
};

}  // namespace nestgpu
